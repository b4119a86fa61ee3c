// The learned projection guess: a 4-level U-Net (16/32/48/64 channels, 3x3
// convolutions) in BF16 on the MFMA units with FP32 accumulation, with its
// own backward kernels and Adam, so that it is trained on the GPU.
//
// Layout. Each level holds its activations channels-last (NHWC: node (x,y),
// channel c at (y*nx + x)*C + c) in BF16; gradients are FP32 in the same
// layout. Level l+1 keeps every other node of level l ((n + 1) / 2 nodes),
// the down-sampling layers are stride-2 3x3 convolutions centred on the kept
// nodes, and the up-sampling is the vertex-centred bilinear interpolation
// (the multigrid's prolongation), clamped at an even-sized edge.
//
// Convolution as implicit GEMM. A workgroup of four wavefronts computes four
// output rows of 64 nodes (the GEMM's M, four 16-node segments per
// wavefront) for every output channel (N, in 16-wide MFMA blocks). K runs
// over 16-channel groups of the input, each group staged once into LDS as
// the input tile the 64 x 4 outputs read (with its halo), then over the
// nine taps: K index = (group * 9 + tap) * 16 + channel % 16, so one K chunk
// of 16 is one tap of one channel group and a lane's four k values are four
// consecutive channels (one 8-byte LDS read). The
// MFMA operand layout (v_mfma_f32_16x16x16_bf16): A row = lane % 16,
// A/B k = 4 (lane / 16) + i, B column = lane % 16; D row = 4 (lane / 16) + i,
// D column = lane % 16.
//   forward:         out(p)  = sum_t W_t in(s p + o_t)             (s = 1, 2)
//   backward-data:   gin(p)  = sum_t W_t^T dpre((p - o_t) / s)       (the same
//                    kernel on the transposed weights, gathering only where
//                    p - o_t is a multiple of s)
//   backward-weight: dW_t    = sum_p dpre(p) in(s p + o_t)^T          (M = output
//                    channels, N = input channels, K = output nodes)
#include <cmath>
#include <cstring>
#include <random>
#include "unet.hh"
#include "common.hh"
#include "../../include/incrmt.h"

namespace incrmt {

namespace {

typedef short shortx4 __attribute__((ext_vector_type(4)));
typedef float floatx4 __attribute__((ext_vector_type(4)));

constexpr int kThreads = 256;  // four wavefronts

struct ConvArgs {
    const bf16_t *in1, *in2;
    int c1, c2;          // channels of the two inputs (multiples of 4)
    int groups;          // 16-channel groups of the input: K = 9 * 16 * groups
    int inx, iny;
    const bf16_t *w;     // [rows][K]
    const float *bias;   // or null
    int cout;            // output channels
    int onx, ony;
    int mode;            // 0: stride 1, 1: stride 2, 2: stride-1 transpose, 3: stride-2 transpose
    int relu;
    bf16_t *out_bf;      // NHWC, cout channels
    float *out_f;        // channels [0, split), NHWC with split channels
    float *out_f2;       // channels [split, cout), NHWC with cout - split channels
    int split, accum;
};

__host__ __device__ inline int kidx(int tap, int c) { return ((c >> 4) * 9 + tap) * 16 + (c & 15); }

// The input tile a workgroup's outputs [ox0, ox0+64) x [oy0, oy0+4) read
// (one more than needed on each side keeps the bounds simple): its first
// point and extent per direction for each mode, and the tile coordinates
// (column, row) output column x and row y of the workgroup read at tap
// (dx, dy). Points outside the input are staged as zeros, so the taps read
// the tile without bounds checks; only the stride-2 transpose needs a
// parity mask.
template <int MODE> struct Tile {
    static constexpr int W = MODE == 1 ? 129 : MODE == 3 ? 34 : 66;
    static constexpr int H = MODE == 1 ? 9 : MODE == 3 ? 4 : 6;
    __device__ static int x0(int o0) { return MODE == 1 ? 2 * o0 - 1 : MODE == 3 ? (o0 >> 1) - 1 : o0 - 1; }
    // tile index of output offset u (0..63 or 0..3) at tap offset d
    __device__ static constexpr int at(int u, int d) {
        return MODE == 0 ? u + d + 1 : MODE == 1 ? 2 * u + d + 1 : MODE == 2 ? u - d + 1 : ((u - d) >> 1) + 1;
    }
};

template <int NB, int MODE>
__global__ __launch_bounds__(kThreads) void k_conv(ConvArgs a) {
    typedef Tile<MODE> T;
    __shared__ __align__(16) bf16_t tile[T::H * T::W * 16];
    __shared__ __align__(16) bf16_t wl[NB * 16 * 144];  // the weights of one channel group
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int ox0 = blockIdx.x * 64, oy0 = blockIdx.y * 4, oy = oy0 + wave;
    const int pl = lane & 15, kg = lane >> 4;
    const int cin = a.c1 + a.c2, kp = a.groups * 144;
    const int tx0 = T::x0(ox0), ty0 = T::x0(oy0);
    floatx4 acc[4][NB];
#pragma unroll
    for (int mt = 0; mt < 4; mt++)
#pragma unroll
        for (int nb = 0; nb < NB; nb++) acc[mt][nb] = floatx4{0.f, 0.f, 0.f, 0.f};
    for (int g = 0; g < a.groups; g++) {
        __syncthreads();
        // the group's weights: NB*16 rows of 144 (9 taps x 16 channels)
        for (int q0 = threadIdx.x; q0 < NB * 16 * 36; q0 += 4 * kThreads) {
            shortx4 v[4];
#pragma unroll
            for (int u = 0; u < 4; u++) {
                const int q = q0 + u * kThreads, row = q / 36, k = (q - row * 36) * 4;
                v[u] = q < NB * 16 * 36 ? *reinterpret_cast<const shortx4 *>(a.w + (size_t)row * kp + g * 144 + k)
                                        : shortx4{0, 0, 0, 0};
            }
#pragma unroll
            for (int u = 0; u < 4; u++)
                if (q0 + u * kThreads < NB * 16 * 36) *reinterpret_cast<shortx4 *>(&wl[(q0 + u * kThreads) * 4]) = v[u];
        }
        // the input tile of channel group g (4-channel quads, zero outside
        // the input and past its channels): a thread takes one column quad
        // of every tile row, all rows' loads in flight together
        for (int idx = threadIdx.x; idx < T::W * 4; idx += kThreads) {
            const int col = idx >> 2, c = g * 16 + (idx & 3) * 4, sx = tx0 + col;
            const bool cok = c < cin && sx >= 0 && sx < a.inx;
            const bf16_t *base = c < a.c1 ? a.in1 + c : a.in2 + (c - a.c1);
            const int cstride = c < a.c1 ? a.c1 : a.c2;
            shortx4 v[T::H];
#pragma unroll
            for (int r = 0; r < T::H; r++) {
                const int sy = ty0 + r;
                v[r] = shortx4{0, 0, 0, 0};
                if (cok && sy >= 0 && sy < a.iny)
                    v[r] = *reinterpret_cast<const shortx4 *>(base + ((size_t)sy * a.inx + sx) * cstride);
            }
#pragma unroll
            for (int r = 0; r < T::H; r++) *reinterpret_cast<shortx4 *>(&tile[(r * T::W + col) * 16 + (idx & 3) * 4]) = v[r];
        }
        __syncthreads();
#pragma unroll
        for (int tap = 0; tap < 9; tap++) {
            const int dx = tap % 3 - 1, dy = tap / 3 - 1;
            // stride-2 transpose: a row of the wrong parity reads nothing
            if (MODE == 3 && ((wave - dy) & 1)) continue;
            shortx4 bv[NB];
#pragma unroll
            for (int nb = 0; nb < NB; nb++)
                bv[nb] = *reinterpret_cast<const shortx4 *>(&wl[(nb * 16 + pl) * 144 + tap * 16 + kg * 4]);
            const int row = T::at(wave, dy);
            shortx4 av[4];
#pragma unroll
            for (int mt = 0; mt < 4; mt++) {
                av[mt] = *reinterpret_cast<const shortx4 *>(&tile[(row * T::W + T::at(mt * 16 + pl, dx)) * 16 + kg * 4]);
                if (MODE == 3 && ((pl - dx) & 1)) av[mt] = shortx4{0, 0, 0, 0};
            }
#pragma unroll
            for (int mt = 0; mt < 4; mt++)
#pragma unroll
                for (int nb = 0; nb < NB; nb++)
                    acc[mt][nb] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(av[mt], bv[nb], acc[mt][nb], 0, 0, 0);
        }
    }
    if (MODE <= 1 && a.out_bf && !a.out_f && a.cout % 16 == 0) {
        // BF16 activations: through LDS (this wavefront's quarter of the
        // tile) so that each lane stores four channels of a node, 8 bytes
        __syncthreads();
        bf16_t *st = tile + wave * 64 * 16;
        for (int nb = 0; nb < NB; nb++) {
            const int ch = nb * 16 + pl;
            const float b = a.bias ? a.bias[ch] : 0.f;
#pragma unroll
            for (int mt = 0; mt < 4; mt++)
#pragma unroll
                for (int i = 0; i < 4; i++) {
                    float v = acc[mt][nb][i] + b;
                    if (a.relu) v = fmaxf(v, 0.f);
                    st[(mt * 16 + 4 * kg + i) * 16 + pl] = f2bf(v);
                }
            __builtin_amdgcn_wave_barrier();
#pragma unroll
            for (int r = 0; r < 4; r++) {
                const int node = r * 16 + (lane >> 2), quad = lane & 3, px = ox0 + node;
                const shortx4 v = *reinterpret_cast<const shortx4 *>(&st[node * 16 + quad * 4]);
                if (oy < a.ony && px < a.onx)
                    *reinterpret_cast<shortx4 *>(a.out_bf + ((size_t)oy * a.onx + px) * a.cout + nb * 16 + quad * 4) = v;
            }
            __builtin_amdgcn_wave_barrier();
        }
        return;
    }
    if (oy >= a.ony) return;
#pragma unroll
    for (int mt = 0; mt < 4; mt++)
#pragma unroll
        for (int nb = 0; nb < NB; nb++)
#pragma unroll
            for (int i = 0; i < 4; i++) {
                const int px = ox0 + mt * 16 + 4 * kg + i, ch = nb * 16 + pl;
                if (px >= a.onx || ch >= a.cout) continue;
                float v = acc[mt][nb][i];
                if (a.bias) v += a.bias[ch];
                if (a.relu) v = fmaxf(v, 0.f);
                const size_t p = (size_t)oy * a.onx + px;
                if (a.out_bf) a.out_bf[p * a.cout + ch] = f2bf(v);
                if (a.out_f) {
                    float *d = ch < a.split ? a.out_f + p * a.split + ch
                                            : a.out_f2 + p * (a.cout - a.split) + (ch - a.split);
                    *d = a.accum ? *d + v : v;
                }
            }
}

struct DwArgs {
    const bf16_t *dpre;
    int cd;               // channels of dpre (output channels rounded up to 4)
    const bf16_t *in1, *in2;
    int c1, c2;
    int inx, iny, onx, ony, stride;
    int kp, cout;
    int nco, nci;         // 16-blocks of output and input channels
    long long segs;       // 16-node row segments of the output grid
    int seg_chunk;        // segments per workgroup
    float *dw;            // [cout_p][kp]
    float *db;            // [cout]: the bias gradient (workgroups of input block 0 add it)
};

// dW of one (output-channel block, input-channel block) pair for all nine
// taps over seg_chunk row segments. Per segment a wavefront loads the
// 16-node x 16-channel blocks of dpre and of the tap-shifted input with
// 8-byte channel-quad loads, transposes them through its own LDS slots
// (the MFMA wants four consecutive nodes of one channel per lane), and
// accumulates one MFMA per tap. The four wavefronts' tiles are summed in
// LDS and added to dW. The workgroups of input block 0 also sum dpre over
// their nodes for the bias gradient.
__global__ __launch_bounds__(kThreads) void k_conv_dw(DwArgs a) {
    __shared__ bf16_t ta[4][256], tb[4][256];
    __shared__ float red[4][256];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int cib = blockIdx.x % a.nci, cob = blockIdx.x / a.nci;
    const int pl = lane & 15, kg = lane >> 4;     // MFMA operand lanes
    const int lp = lane >> 2, lq = (lane & 3) * 4;  // loading lanes: node, channel quad
    const int cin = a.c1 + a.c2;
    const int tiles_x = (a.onx + 15) >> 4;
    const int co = cob * 16 + lq, ci = cib * 16 + lq;
    const bool ci_ok = ci < cin;
    const bf16_t *xin = ci < a.c1 ? a.in1 : a.in2;
    const int xc = ci < a.c1 ? a.c1 : a.c2, xo = ci < a.c1 ? ci : ci - a.c1;
    floatx4 acc[9];
#pragma unroll
    for (int t = 0; t < 9; t++) acc[t] = floatx4{0.f, 0.f, 0.f, 0.f};
    float bsum[4] = {0.f, 0.f, 0.f, 0.f};  // bias sums of this lane's channel quad
    const long long s0 = (long long)blockIdx.y * a.seg_chunk;
    long long s1 = s0 + a.seg_chunk;
    if (s1 > a.segs) s1 = a.segs;
    for (long long s = s0 + wave; s < s1; s += 4) {
        const int oy = (int)(s / tiles_x), ox0 = (int)(s % tiles_x) * 16;
        const int px = ox0 + lp;
        // all ten loads of the segment first, then the transposes
        shortx4 v = {0, 0, 0, 0}, x[9];
        if (px < a.onx && co < a.cd) v = *reinterpret_cast<const shortx4 *>(a.dpre + ((size_t)oy * a.onx + px) * a.cd + co);
#pragma unroll
        for (int t = 0; t < 9; t++) {
            const int sx = a.stride * px + (t % 3 - 1), sy = a.stride * oy + (t / 3 - 1);
            x[t] = shortx4{0, 0, 0, 0};
            if (ci_ok && px < a.onx && sx >= 0 && sy >= 0 && sx < a.inx && sy < a.iny)
                x[t] = *reinterpret_cast<const shortx4 *>(xin + ((size_t)sy * a.inx + sx) * xc + xo);
        }
#pragma unroll
        for (int i = 0; i < 4; i++) {
            ta[wave][(lq + i) * 16 + lp] = (bf16_t)v[i];
            bsum[i] += bf2f((bf16_t)v[i]);
        }
        __builtin_amdgcn_wave_barrier();
        const shortx4 av = *reinterpret_cast<const shortx4 *>(&ta[wave][pl * 16 + 4 * kg]);
#pragma unroll
        for (int t = 0; t < 9; t++) {
            __builtin_amdgcn_wave_barrier();
#pragma unroll
            for (int i = 0; i < 4; i++) tb[wave][(lq + i) * 16 + lp] = (bf16_t)x[t][i];
            __builtin_amdgcn_wave_barrier();
            const shortx4 bv = *reinterpret_cast<const shortx4 *>(&tb[wave][pl * 16 + 4 * kg]);
            acc[t] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(av, bv, acc[t], 0, 0, 0);
        }
        __builtin_amdgcn_wave_barrier();
    }
    // D row = output channel 4 kg + i of the block, column = input channel pl
    const int e = threadIdx.x, r = e >> 4, cc = e & 15;
    const int gco = cob * 16 + r, gci = cib * 16 + cc;
    for (int t = 0; t < 9; t++) {
#pragma unroll
        for (int i = 0; i < 4; i++) red[wave][(4 * kg + i) * 16 + pl] = acc[t][i];
        __syncthreads();
        const float v = red[0][e] + red[1][e] + red[2][e] + red[3][e];
        if (gco < a.cout && gci < cin && v != 0.f) atomicAdd(&a.dw[(size_t)gco * a.kp + kidx(t, gci)], v);
        __syncthreads();
    }
    if (cib != 0) return;
    // bias: sum the 16 nodes (lanes lp) of each channel quad, then the waves
#pragma unroll
    for (int i = 0; i < 4; i++) red[wave][(lq + i) * 16 + lp] = bsum[i];
    __syncthreads();
    if (threadIdx.x < 16) {
        float b = 0.f;
        for (int w = 0; w < 4; w++)
            for (int n = 0; n < 16; n++) b += red[w][threadIdx.x * 16 + n];
        const int ch = cob * 16 + threadIdx.x;
        if (ch < a.cout && b != 0.f) atomicAdd(&a.db[ch], b);
    }
}

// dpre = d loss / d (pre-activation) in BF16: the output gradient where the
// ReLU passed (act > 0), zero elsewhere and on the padding channels
__global__ void k_relu_bwd(const bf16_t *act, const float *g, int cout, int cd, long long npix, int relu,
                           bf16_t *dpre) {
    const long long n = npix * cd;
    for (long long q = blockIdx.x * (long long)blockDim.x + threadIdx.x; q < n; q += (long long)gridDim.x * blockDim.x) {
        const long long p = q / cd;
        const int c = (int)(q - p * cd);
        float v = 0.f;
        if (c < cout) {
            v = g[p * cout + c];
            if (relu && !(bf2f(act[p * cout + c]) > 0.f)) v = 0.f;
        }
        dpre[q] = f2bf(v);
    }
}

// vertex-centred bilinear prolongation: fine node x sits at coarse x/2
__device__ inline void parents(int x, int cn, int &x0, int &x1, float &w0, float &w1) {
    if (!(x & 1)) { x0 = x1 = x >> 1; w0 = 1.f; w1 = 0.f; return; }
    x0 = x >> 1; x1 = x0 + 1;
    if (x1 > cn - 1) x1 = cn - 1;
    w0 = w1 = 0.5f;
}
// one thread per fine node and channel quad (8-byte loads and stores)
__global__ void k_upsample(const bf16_t *cin, int cnx, int cny, int C, bf16_t *fout, int fnx, int fny) {
    const int Q = C >> 2;
    const long long n = (long long)fnx * fny * Q;
    for (long long t = blockIdx.x * (long long)blockDim.x + threadIdx.x; t < n; t += (long long)gridDim.x * blockDim.x) {
        const long long p = t / Q;
        const int c = (int)(t - p * Q) * 4, x = (int)(p % fnx), y = (int)(p / fnx);
        int x0, x1, y0, y1;
        float wx0, wx1, wy0, wy1;
        parents(x, cnx, x0, x1, wx0, wx1);
        parents(y, cny, y0, y1, wy0, wy1);
        const shortx4 a00 = *reinterpret_cast<const shortx4 *>(cin + ((size_t)y0 * cnx + x0) * C + c);
        const shortx4 a10 = *reinterpret_cast<const shortx4 *>(cin + ((size_t)y0 * cnx + x1) * C + c);
        const shortx4 a01 = *reinterpret_cast<const shortx4 *>(cin + ((size_t)y1 * cnx + x0) * C + c);
        const shortx4 a11 = *reinterpret_cast<const shortx4 *>(cin + ((size_t)y1 * cnx + x1) * C + c);
        shortx4 o;
#pragma unroll
        for (int i = 0; i < 4; i++) {
            const float v = wy0 * (wx0 * bf2f((bf16_t)a00[i]) + wx1 * bf2f((bf16_t)a10[i])) +
                            wy1 * (wx0 * bf2f((bf16_t)a01[i]) + wx1 * bf2f((bf16_t)a11[i]));
            o[i] = (short)f2bf(v);
        }
        *reinterpret_cast<shortx4 *>(fout + p * C + c) = o;
    }
}
// its adjoint, gathered per coarse node: the fine nodes 2X-1 .. 2X+1 of
// each direction with the weights they give X
__device__ inline float weight_to(int x, int X, int cn) {
    int x0, x1;
    float w0, w1;
    parents(x, cn, x0, x1, w0, w1);
    return (x0 == X ? w0 : 0.f) + (x1 == X ? w1 : 0.f);
}
__global__ void k_upsample_bwd(const float *gf, int fnx, int fny, int C, float *gc, int cnx, int cny) {
    const long long n = (long long)cnx * cny * C;
    for (long long q = blockIdx.x * (long long)blockDim.x + threadIdx.x; q < n; q += (long long)gridDim.x * blockDim.x) {
        const long long p = q / C;
        const int c = (int)(q - p * C), X = (int)(p % cnx), Y = (int)(p / cnx);
        float s = 0.f;
        for (int y = 2 * Y - 1; y <= 2 * Y + 1; y++) {
            if (y < 0 || y >= fny) continue;
            const float wy = weight_to(y, Y, cny);
            if (wy == 0.f) continue;
            for (int x = 2 * X - 1; x <= 2 * X + 1; x++) {
                if (x < 0 || x >= fnx) continue;
                const float wx = weight_to(x, X, cnx);
                if (wx != 0.f) s += wx * wy * gf[((size_t)y * fnx + x) * C + c];
            }
        }
        gc[q] = s;
    }
}

// loss = 0.5 mean((out - t)^2); gout = (out - t) / N; deterministic
// two-level sum of the loss
__global__ __launch_bounds__(kThreads) void k_loss(const float *out, const float *t, long long n, float *gout,
                                                   double *part) {
    __shared__ double s[kThreads];
    double a = 0;
    const double inv = 1.0 / (double)n;
    for (long long p = blockIdx.x * (long long)blockDim.x + threadIdx.x; p < n; p += (long long)gridDim.x * blockDim.x) {
        const float d = out[p] - t[p];
        gout[p] = (float)(d * inv);
        a += 0.5 * (double)d * d * inv;
    }
    s[threadIdx.x] = a;
    __syncthreads();
    for (int w = kThreads / 2; w > 0; w >>= 1) {
        if (threadIdx.x < w) s[threadIdx.x] += s[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) part[1 + blockIdx.x] = s[0];
}
__global__ void k_loss_final(double *part, int nb) {
    if (threadIdx.x == 0 && blockIdx.x == 0) {
        double a = 0;
        for (int b = 0; b < nb; b++) a += part[1 + b];
        part[0] = a;
    }
}

__global__ void k_adam(float *w, const float *g, float *m, float *v, long long n, float lr, float c1, float c2) {
    for (long long q = blockIdx.x * (long long)blockDim.x + threadIdx.x; q < n; q += (long long)gridDim.x * blockDim.x) {
        const float gq = g[q];
        const float mq = 0.9f * m[q] + 0.1f * gq, vq = 0.999f * v[q] + 0.001f * gq * gq;
        m[q] = mq;
        v[q] = vq;
        w[q] -= lr * (mq * c1) / (sqrtf(vq * c2) + 1e-8f);
    }
}

__global__ void k_to_bf16(const float *a, bf16_t *b, long long n) {
    for (long long q = blockIdx.x * (long long)blockDim.x + threadIdx.x; q < n; q += (long long)gridDim.x * blockDim.x)
        b[q] = f2bf(a[q]);
}
// transposed weights of one layer for the backward-data pass, whose input
// channels are the layer's output channels: wt[ci][kidx(tap, co)] =
// w[co][kidx(tap, ci)] (the transposed convolution flips the offsets, which
// the backward-data modes do through their gather)
__global__ void k_build_wt(const float *w, int kp, int cin, int cout, int rows, int kpt, bf16_t *wt) {
    const long long n = (long long)rows * kpt;
    for (long long q = blockIdx.x * (long long)blockDim.x + threadIdx.x; q < n; q += (long long)gridDim.x * blockDim.x) {
        const int ci = (int)(q / kpt), k = (int)(q - (long long)ci * kpt);
        const int grp = k / 144, rem = k - grp * 144, tap = rem >> 4, co = grp * 16 + (rem & 15);
        float v = 0.f;
        if (ci < cin && co < cout) v = w[(size_t)co * kp + kidx(tap, ci)];
        wt[q] = f2bf(v);
    }
}
__global__ void k_fill_f32(float *p, long long n, float v) {
    for (long long q = blockIdx.x * (long long)blockDim.x + threadIdx.x; q < n; q += (long long)gridDim.x * blockDim.x)
        p[q] = v;
}
__global__ void k_pack_input(const float *in3, long long npix, bf16_t *in4) {
    for (long long q = blockIdx.x * (long long)blockDim.x + threadIdx.x; q < npix * 4;
         q += (long long)gridDim.x * blockDim.x) {
        const long long p = q >> 2;
        const int c = (int)(q & 3);
        in4[q] = f2bf(c < 3 ? in3[p * 3 + c] : 0.f);
    }
}

inline int blocks_for(long long n) {
    long long b = (n + kThreads - 1) / kThreads;
    return (int)(b > 8192 ? 8192 : (b < 1 ? 1 : b));
}
inline int cd_of(int cout) { return (cout + 3) / 4 * 4; }

}  // namespace

#define UCHECK() HIP_CHECK(hipGetLastError())

int UNet::read_loss(double *loss) {
    HIP_CHECK(hipMemcpyAsync(h_loss_, d_loss_, sizeof(double), hipMemcpyDeviceToHost, s_));
    HIP_CHECK(hipStreamSynchronize(s_));
    *loss = *h_loss_;
    return 0;
}

UNet::~UNet() {
    for (void *p : allocs_) (void)hipFree(p);
    if (h_pack_) (void)hipHostFree(h_pack_);
    if (h_loss_) (void)hipHostFree(h_loss_);
}

int UNet::alloc(void **p, size_t bytes) {
    HIP_CHECK(hipMalloc(p, bytes));
    HIP_CHECK(memset_sync(*p, 0, bytes, s_));
    allocs_.push_back(*p);
    return 0;
}

int UNet::init(int nx, int ny, unsigned seed, hipStream_t s) {
    if (nx < 8 || ny < 8) return INCRMT_ERR_ARG;
    s_ = s;
    nx_[0] = nx; ny_[0] = ny;
    for (int l = 1; l < kUnetLevels; l++) { nx_[l] = (nx_[l - 1] + 1) / 2; ny_[l] = (ny_[l - 1] + 1) / 2; }
    const int C0 = kUnetCh[0], C1 = kUnetCh[1], C2 = kUnetCh[2], C3 = kUnetCh[3];
    // cin1, cin2, cout, stride, lin, lout, relu
    const int spec[kUnetLayers][7] = {
        {kUnetIn, 0, C0, 1, 0, 0, 1}, {C0, 0, C0, 1, 0, 0, 1},  // E0a, E0b
        {C0, 0, C1, 2, 0, 1, 1},      {C1, 0, C1, 1, 1, 1, 1},  // D1, E1
        {C1, 0, C2, 2, 1, 2, 1},      {C2, 0, C2, 1, 2, 2, 1},  // D2, E2
        {C2, 0, C3, 2, 2, 3, 1},      {C3, 0, C3, 1, 3, 3, 1},  // D3, B3
        {C3, C2, C2, 1, 2, 2, 1},     {C2, C1, C1, 1, 1, 1, 1}, // U2, U1
        {C1, C0, C0, 1, 0, 0, 1},     {C0, 0, 1, 1, 0, 0, 0},   // U0, Out
    };
    size_t off = 0, packed = 0, wt = 0;
    for (int l = 0; l < kUnetLayers; l++) {
        ConvSpec &c = L_[l];
        c.cin1 = spec[l][0]; c.cin2 = spec[l][1]; c.cout = spec[l][2]; c.stride = spec[l][3];
        c.lin = spec[l][4]; c.lout = spec[l][5]; c.relu = spec[l][6] != 0;
        c.groups = (c.cin1 + c.cin2 + 15) / 16;
        c.kp = 144 * c.groups;
        c.cout_p = (c.cout + 15) / 16 * 16;
        c.w_off = off; off += (size_t)c.cout_p * c.kp;
        c.b_off = off; off += (size_t)c.cout_p;
        packed += (size_t)c.cout * 9 * (c.cin1 + c.cin2) + c.cout;
        // transposed weights: rows = input channels rounded to 16, K = 144
        // per 16-channel group of the output channels
        wt_off_[l] = wt;
        wt += (size_t)(16 * c.groups) * (144 * ((c.cout + 15) / 16));
    }
    nmaster_ = off;
    npacked_ = packed;
    if (int rc = alloc((void **)&w32_, off * 4)) return rc;
    if (int rc = alloc((void **)&g32_, off * 4)) return rc;
    if (int rc = alloc((void **)&m_, off * 4)) return rc;
    if (int rc = alloc((void **)&v_, off * 4)) return rc;
    if (int rc = alloc((void **)&w16_, off * 2)) return rc;
    if (int rc = alloc((void **)&wt16_, wt * 2)) return rc;
    auto npix = [&](int l) { return (size_t)nx_[l] * ny_[l]; };
    if (int rc = alloc((void **)&in_, npix(0) * kUnetIn * 2)) return rc;
    size_t dpre = 0;
    for (int l = 0; l < kUnetLayers; l++) {
        const ConvSpec &c = L_[l];
        if (l != kOut) {
            if (int rc = alloc((void **)&act_[l], npix(c.lout) * c.cout * 2)) return rc;
        }
        if (int rc = alloc((void **)&gact_[l], npix(c.lout) * c.cout * 4)) return rc;
        if (npix(c.lout) * cd_of(c.cout) > dpre) dpre = npix(c.lout) * cd_of(c.cout);
    }
    gout_ = gact_[kOut];
    if (int rc = alloc((void **)&dpre_, dpre * 2)) return rc;
    for (int u = 0; u < 3; u++) {  // inputs of kU2 (level 2), kU1 (1), kU0 (0)
        const int lev = 2 - u, C = kUnetCh[lev + 1];
        if (int rc = alloc((void **)&up_[u], npix(lev) * C * 2)) return rc;
        if (int rc = alloc((void **)&gup_[u], npix(lev) * C * 4)) return rc;
    }
    if (int rc = alloc((void **)&out_, npix(0) * 4)) return rc;
    if (int rc = alloc((void **)&d_loss_, (1 + 1024) * sizeof(double))) return rc;
    HIP_CHECK(hipHostMalloc((void **)&h_pack_, (off > packed ? off : packed) * 4, hipHostMallocDefault));
    HIP_CHECK(hipHostMalloc((void **)&h_loss_, sizeof(double), hipHostMallocDefault));
    // He-normal weights (std sqrt(2 / fan-in)); the output layer starts
    // small so the first predictions are near zero; zero biases
    std::mt19937_64 rng(seed);
    std::vector<float> h(off, 0.f);
    for (int l = 0; l < kUnetLayers; l++) {
        const ConvSpec &c = L_[l];
        const int cin = c.cin1 + c.cin2;
        std::normal_distribution<double> nd(0.0, std::sqrt(2.0 / (9.0 * cin)) * (l == kOut ? 0.1 : 1.0));
        for (int co = 0; co < c.cout; co++)
            for (int t = 0; t < 9; t++)
                for (int ci = 0; ci < cin; ci++) {
                    h[c.w_off + (size_t)co * c.kp + kidx(t, ci)] = (float)nd(rng);
                }
    }
    HIP_CHECK(hipMemcpy(w32_, h.data(), off * 4, hipMemcpyHostToDevice));
    w_dirty_ = true;
    return 0;
}

int UNet::refresh_weights() {
    if (!w_dirty_) return 0;
    k_to_bf16<<<blocks_for(nmaster_), kThreads, 0, s_>>>(w32_, w16_, (long long)nmaster_);
    UCHECK();
    for (int l = 0; l < kUnetLayers; l++) {
        const ConvSpec &c = L_[l];
        const int cin = c.cin1 + c.cin2, rows = 16 * c.groups, kpt = 144 * ((c.cout + 15) / 16);
        k_build_wt<<<blocks_for((long long)rows * kpt), kThreads, 0, s_>>>(w32_ + c.w_off, c.kp, cin, c.cout, rows,
                                                                           kpt, wt16_ + wt_off_[l]);
        UCHECK();
    }
    w_dirty_ = false;
    return 0;
}

template <int MODE>
static int launch_conv_m(int nb, const ConvArgs &a, hipStream_t s) {
    const dim3 grid((a.onx + 63) / 64, (a.ony + 3) / 4);
    switch (nb) {
        case 1: k_conv<1, MODE><<<grid, kThreads, 0, s>>>(a); break;
        case 2: k_conv<2, MODE><<<grid, kThreads, 0, s>>>(a); break;
        case 3: k_conv<3, MODE><<<grid, kThreads, 0, s>>>(a); break;
        case 4: k_conv<4, MODE><<<grid, kThreads, 0, s>>>(a); break;
        case 5: k_conv<5, MODE><<<grid, kThreads, 0, s>>>(a); break;
        case 7: k_conv<7, MODE><<<grid, kThreads, 0, s>>>(a); break;
        default: return INCRMT_ERR_ARG;
    }
    HIP_CHECK(hipGetLastError());
    return 0;
}
static int launch_conv(int nb, const ConvArgs &a, hipStream_t s) {
    switch (a.mode) {
        case 0: return launch_conv_m<0>(nb, a, s);
        case 1: return launch_conv_m<1>(nb, a, s);
        case 2: return launch_conv_m<2>(nb, a, s);
        default: return launch_conv_m<3>(nb, a, s);
    }
}

int UNet::conv_forward(int l, const bf16_t *in1, const bf16_t *in2, bf16_t *out_bf, float *out_f) {
    const ConvSpec &c = L_[l];
    ConvArgs a{};
    a.in1 = in1; a.in2 = in2; a.c1 = c.cin1; a.c2 = c.cin2;
    a.groups = c.groups;
    a.inx = nx_[c.lin]; a.iny = ny_[c.lin];
    a.w = w16_ + c.w_off; a.bias = w32_ + c.b_off;
    a.cout = c.cout; a.onx = nx_[c.lout]; a.ony = ny_[c.lout];
    a.mode = c.stride == 2 ? 1 : 0; a.relu = c.relu;
    a.out_bf = out_bf; a.out_f = out_f; a.out_f2 = nullptr; a.split = c.cout; a.accum = 0;
    return launch_conv(c.cout_p / 16, a, s_);
}

// Backward through layer l: dpre from the output gradient (and the ReLU
// mask), then dW and db (accumulated into g32_), then, when gin1 is given,
// the input gradients (gin1 for the first input, gin2 for the skip input of
// a decoder layer), added to what the buffers hold when accum is set.
int UNet::conv_backward(int l, const bf16_t *in1, const bf16_t *in2, const bf16_t *act_out, const float *gout,
                        float *gin1, float *gin2, bool accum) {
    const ConvSpec &c = L_[l];
    const long long npo = (long long)nx_[c.lout] * ny_[c.lout];
    const int cd = cd_of(c.cout);
    k_relu_bwd<<<blocks_for(npo * cd), kThreads, 0, s_>>>(act_out, gout, c.cout, cd, npo, c.relu ? 1 : 0, dpre_);
    UCHECK();
    // weights
    DwArgs d{};
    d.dpre = dpre_; d.cd = cd; d.in1 = in1; d.in2 = in2; d.c1 = c.cin1; d.c2 = c.cin2;
    d.inx = nx_[c.lin]; d.iny = ny_[c.lin]; d.onx = nx_[c.lout]; d.ony = ny_[c.lout]; d.stride = c.stride;
    d.kp = c.kp; d.cout = c.cout; d.db = g32_ + c.b_off;
    d.nco = c.cout_p / 16; d.nci = (c.cin1 + c.cin2 + 15) / 16;
    d.segs = (long long)((d.onx + 15) / 16) * d.ony;
    // about 1024 workgroups in all, each over at least 16 segments
    const int tiles = d.nco * d.nci;
    long long want = (1024 + tiles - 1) / tiles;
    long long chunk = (d.segs + want - 1) / want;
    d.seg_chunk = (int)(chunk < 16 ? 16 : chunk);
    d.dw = g32_ + c.w_off;
    const long long chunks = (d.segs + d.seg_chunk - 1) / d.seg_chunk;
    k_conv_dw<<<dim3(tiles, (unsigned)chunks), kThreads, 0, s_>>>(d);
    UCHECK();
    if (!gin1) return 0;
    // input gradients: the transposed convolution
    const int cin = c.cin1 + c.cin2;
    ConvArgs a{};
    a.in1 = dpre_; a.in2 = nullptr; a.c1 = cd; a.c2 = 0;
    a.groups = (cd + 15) / 16;
    a.inx = nx_[c.lout]; a.iny = ny_[c.lout];
    a.w = wt16_ + wt_off_[l]; a.bias = nullptr;
    a.cout = cin; a.onx = nx_[c.lin]; a.ony = ny_[c.lin];
    a.mode = c.stride == 2 ? 3 : 2; a.relu = 0;
    a.out_bf = nullptr; a.out_f = gin1; a.out_f2 = gin2; a.split = c.cin1;
    a.accum = accum ? 1 : 0;
    return launch_conv((cin + 15) / 16, a, s_);
}

int UNet::forward(const bf16_t *d_in, float *d_out) {
    if (int rc = refresh_weights()) return rc;
    int rc;
    if ((rc = conv_forward(kE0a, d_in, nullptr, act_[kE0a], nullptr)) ||
        (rc = conv_forward(kE0b, act_[kE0a], nullptr, act_[kE0b], nullptr)) ||
        (rc = conv_forward(kD1, act_[kE0b], nullptr, act_[kD1], nullptr)) ||
        (rc = conv_forward(kE1, act_[kD1], nullptr, act_[kE1], nullptr)) ||
        (rc = conv_forward(kD2, act_[kE1], nullptr, act_[kD2], nullptr)) ||
        (rc = conv_forward(kE2, act_[kD2], nullptr, act_[kE2], nullptr)) ||
        (rc = conv_forward(kD3, act_[kE2], nullptr, act_[kD3], nullptr)) ||
        (rc = conv_forward(kB3, act_[kD3], nullptr, act_[kB3], nullptr)))
        return rc;
    const int lev[3] = {2, 1, 0};
    const int from[3] = {kB3, kU2, kU1}, layer[3] = {kU2, kU1, kU0}, skip[3] = {kE2, kE1, kE0b};
    for (int u = 0; u < 3; u++) {
        const int lf = lev[u], lc = lf + 1, C = kUnetCh[lc];
        const long long n = (long long)nx_[lf] * ny_[lf] * (C / 4);
        k_upsample<<<blocks_for(n), kThreads, 0, s_>>>(act_[from[u]], nx_[lc], ny_[lc], C, up_[u], nx_[lf], ny_[lf]);
        UCHECK();
        if ((rc = conv_forward(layer[u], up_[u], act_[skip[u]], act_[layer[u]], nullptr))) return rc;
    }
    if ((rc = conv_forward(kOut, act_[kU0], nullptr, nullptr, out_))) return rc;
    if (d_out && d_out != out_)
        HIP_CHECK(hipMemcpyAsync(d_out, out_, (size_t)nx_[0] * ny_[0] * 4, hipMemcpyDeviceToDevice, s_));
    last_in_ = d_in;
    return 0;
}

int UNet::backward(const float *d_target, double *loss) {
    if (!last_in_) return INCRMT_ERR_ARG;
    const long long n0 = (long long)nx_[0] * ny_[0];
    const int nb = 1024;
    k_loss<<<nb, kThreads, 0, s_>>>(out_, d_target, n0, gout_, d_loss_);
    UCHECK();
    k_loss_final<<<1, 64, 0, s_>>>(d_loss_, nb);
    UCHECK();
    if (int rc = backward_grad(gout_)) return rc;
    if (loss) {
        HIP_CHECK(hipMemcpyAsync(h_loss_, d_loss_, sizeof(double), hipMemcpyDeviceToHost, s_));
        HIP_CHECK(hipStreamSynchronize(s_));
        *loss = *h_loss_;
    }
    return 0;
}

int UNet::backward_grad(const float *gout) {
    if (!last_in_) return INCRMT_ERR_ARG;
    k_fill_f32<<<blocks_for(nmaster_), kThreads, 0, s_>>>(g32_, (long long)nmaster_, 0.f);
    UCHECK();
    int rc;
    // decoder: Out, then U0 (up0 | e0), U1, U2, each followed by the adjoint
    // of the up-sampling that fed it
    if ((rc = conv_backward(kOut, act_[kU0], nullptr, nullptr, gout, gact_[kU0], nullptr, false))) return rc;
    const int lev[3] = {0, 1, 2};
    const int layer[3] = {kU0, kU1, kU2}, from[3] = {kU1, kU2, kB3}, skip[3] = {kE0b, kE1, kE2};
    for (int u = 0; u < 3; u++) {
        const int k = 2 - u;  // up_[k] feeds layer[u]
        if ((rc = conv_backward(layer[u], up_[k], act_[skip[u]], act_[layer[u]], gact_[layer[u]], gup_[k],
                                gact_[skip[u]], false)))
            return rc;
        const int lf = lev[u], lc = lf + 1, C = kUnetCh[lc];
        const long long n = (long long)nx_[lc] * ny_[lc] * C;
        k_upsample_bwd<<<blocks_for(n), kThreads, 0, s_>>>(gup_[k], nx_[lf], ny_[lf], C, gact_[from[u]], nx_[lc],
                                                            ny_[lc]);
        UCHECK();
    }
    // encoder, from the bottom; a down-sampling layer adds its input
    // gradient to the skip connection's, which the decoder wrote first
    struct Step { int l, in; bool accum; };
    const Step steps[] = {{kB3, kD3, false}, {kD3, kE2, true}, {kE2, kD2, false}, {kD2, kE1, true},
                          {kE1, kD1, false}, {kD1, kE0b, true}, {kE0b, kE0a, false}};
    for (const Step &st : steps)
        if ((rc = conv_backward(st.l, act_[st.in], nullptr, act_[st.l], gact_[st.l], gact_[st.in], nullptr, st.accum)))
            return rc;
    return conv_backward(kE0a, last_in_, nullptr, act_[kE0a], gact_[kE0a], nullptr, nullptr, false);
}

int UNet::adam(float lr) {
    adam_t_++;
    const double c1 = 1.0 / (1.0 - std::pow(0.9, (double)adam_t_)), c2 = 1.0 / (1.0 - std::pow(0.999, (double)adam_t_));
    k_adam<<<blocks_for(nmaster_), kThreads, 0, s_>>>(w32_, g32_, m_, v_, (long long)nmaster_, lr, (float)c1, (float)c2);
    UCHECK();
    w_dirty_ = true;
    return 0;
}

// packed order: per layer, weights [cout][tap][cin] then bias [cout]
static void pack(const ConvSpec *L, const float *master, float *packed, bool to_packed) {
    size_t q = 0;
    for (int l = 0; l < kUnetLayers; l++) {
        const ConvSpec &c = L[l];
        const int cin = c.cin1 + c.cin2;
        for (int co = 0; co < c.cout; co++)
            for (int t = 0; t < 9; t++)
                for (int ci = 0; ci < cin; ci++, q++) {
                    const size_t m = c.w_off + (size_t)co * c.kp + kidx(t, ci);
                    if (to_packed) packed[q] = master[m];
                    else const_cast<float *>(master)[m] = packed[q];
                }
        for (int co = 0; co < c.cout; co++, q++) {
            if (to_packed) packed[q] = master[c.b_off + co];
            else const_cast<float *>(master)[c.b_off + co] = packed[q];
        }
    }
}

int UNet::get_params(float *host) const {
    std::vector<float> m(nmaster_);
    HIP_CHECK(hipStreamSynchronize(s_));
    HIP_CHECK(hipMemcpy(m.data(), w32_, nmaster_ * 4, hipMemcpyDeviceToHost));
    pack(L_, m.data(), host, true);
    return 0;
}

int UNet::set_params(const float *host) {
    std::vector<float> m(nmaster_, 0.f);
    pack(L_, m.data(), const_cast<float *>(host), false);
    // the padding input channel of the first layer stays zero
    HIP_CHECK(hipStreamSynchronize(s_));
    HIP_CHECK(hipMemcpy(w32_, m.data(), nmaster_ * 4, hipMemcpyHostToDevice));
    w_dirty_ = true;
    return 0;
}

int UNet::get_grads(float *host) const {
    std::vector<float> m(nmaster_);
    HIP_CHECK(hipStreamSynchronize(s_));
    HIP_CHECK(hipMemcpy(m.data(), g32_, nmaster_ * 4, hipMemcpyDeviceToHost));
    pack(L_, m.data(), host, true);
    return 0;
}

int UNet::forward_host(const float *in3, float *out) {
    const long long n0 = (long long)nx_[0] * ny_[0];
    float *tmp = nullptr;
    HIP_CHECK(hipMalloc((void **)&tmp, n0 * 3 * 4));
    HIP_CHECK(hipMemcpy(tmp, in3, n0 * 3 * 4, hipMemcpyHostToDevice));
    k_pack_input<<<blocks_for(n0 * 4), kThreads, 0, s_>>>(tmp, n0, in_);
    int rc = hipGetLastError() == hipSuccess ? forward(in_, out_) : INCRMT_ERR_HIP;
    if (!rc && out) {
        HIP_CHECK(hipStreamSynchronize(s_));
        HIP_CHECK(hipMemcpy(out, out_, n0 * 4, hipMemcpyDeviceToHost));
    }
    HIP_CHECK(hipStreamSynchronize(s_));
    (void)hipFree(tmp);
    return rc;
}

int UNet::backward_host(const float *in3, const float *target, double *loss, float *grads) {
    const long long n0 = (long long)nx_[0] * ny_[0];
    if (int rc = forward_host(in3, nullptr)) return rc;
    float *t = nullptr;
    HIP_CHECK(hipMalloc((void **)&t, n0 * 4));
    HIP_CHECK(hipMemcpy(t, target, n0 * 4, hipMemcpyHostToDevice));
    int rc = backward(t, loss);
    HIP_CHECK(hipStreamSynchronize(s_));
    (void)hipFree(t);
    if (!rc && grads) rc = get_grads(grads);
    return rc;
}

}  // namespace incrmt

// ---------------------------------------------------------------- C ABI
struct incrmt_unet {
    incrmt::UNet net;
    hipStream_t s = nullptr;
    ~incrmt_unet() {
        if (s) (void)hipStreamDestroy(s);
    }
};

extern "C" {

int incrmt_unet_create(int nx, int ny, unsigned seed, incrmt_unet **out) {
    if (!out || nx < 8 || ny < 8) return incrmt::set_error(INCRMT_ERR_ARG, "bad network size");
    incrmt_unet *h = new incrmt_unet;
    hipError_t e = hipStreamCreateWithFlags(&h->s, hipStreamNonBlocking);
    if (e != hipSuccess) {
        delete h;
        return incrmt::hip_fail(e, "hipStreamCreateWithFlags", __FILE__, __LINE__);
    }
    if (int rc = h->net.init(nx, ny, seed, h->s)) {
        delete h;
        return rc;
    }
    *out = h;
    return 0;
}

int incrmt_unet_nparams(const incrmt_unet *h, long long *n) {
    if (!h || !n) return incrmt::set_error(INCRMT_ERR_ARG, "null argument");
    *n = (long long)h->net.nparams();
    return 0;
}

int incrmt_unet_get_params(incrmt_unet *h, float *params) {
    if (!h || !params) return incrmt::set_error(INCRMT_ERR_ARG, "null argument");
    return h->net.get_params(params);
}

int incrmt_unet_set_params(incrmt_unet *h, const float *params) {
    if (!h || !params) return incrmt::set_error(INCRMT_ERR_ARG, "null argument");
    return h->net.set_params(params);
}

int incrmt_unet_forward(incrmt_unet *h, const float *in3, float *out) {
    if (!h || !in3 || !out) return incrmt::set_error(INCRMT_ERR_ARG, "null argument");
    return h->net.forward_host(in3, out);
}

int incrmt_unet_gradient(incrmt_unet *h, const float *in3, const float *target, double *loss, float *grads) {
    if (!h || !in3 || !target) return incrmt::set_error(INCRMT_ERR_ARG, "null argument");
    return h->net.backward_host(in3, target, loss, grads);
}

int incrmt_unet_train_step(incrmt_unet *h, const float *in3, const float *target, float lr, double *loss) {
    if (!h || !in3 || !target || !(lr > 0)) return incrmt::set_error(INCRMT_ERR_ARG, "bad argument");
    if (int rc = h->net.backward_host(in3, target, loss, nullptr)) return rc;
    if (int rc = h->net.adam(lr)) return rc;
    const hipError_t e = hipStreamSynchronize(h->s);
    return e == hipSuccess ? 0 : incrmt::hip_fail(e, "hipStreamSynchronize", __FILE__, __LINE__);
}

void incrmt_unet_destroy(incrmt_unet *h) { delete h; }

}  // extern "C"
