// unet.hip - U-Net image-domain preconditioner for LSRTM (include/erm_unet.h).
//
// Convolutions are implicit GEMMs on gfx950's BF16 MFMA with FP32
// accumulation (v_mfma_f32_16x16x32_bf16).  For a 64-lane wavefront, lane l
// holds row m = l % 16 of A and column n = l % 16 of B, each as 8 k-values of
// lane group g = l / 16; the instruction sums A and B elements over the same
// (g, j), so the kernels are free to choose which logical k each (g, j)
// carries, as long as A and B agree.  D[m][n]: n = l % 16, m = 4*(l/16) + j.
//   forward        Y[co][p]  = sum_{ci,t} W[co][ci][t] X[ci][p+o(t)]   (M=Cout, N=pixels, K=Cin*9)
//   backward data  dX[ci][p] = sum_{co,t} W[co][ci][8-t] dY[co][p+o(t)] (M=Cin,  N=pixels, K=Cout*9)
//   backward wts   dW[co][ci][t] = sum_p dY[co][p] X[ci][p+o(t)]        (M=Cout, N=Cin*9,  K=pixels, split-K)
// with o(t) = (t/3-1, t%3-1) and zero padding; dY is masked by the ReLU of
// the layer output where the layer has one.  Operands are staged in LDS as
// bf16 (round to nearest even) so each lane reads its 8 k-values as one
// 16-byte word.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <random>
#include <string>
#include <vector>

#include "../../include/erm.h"
#include "../../include/erm_unet.h"

namespace {

thread_local std::string g_uerr;

typedef short s16x8 __attribute__((ext_vector_type(8)));
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

// one 16x16x32 MFMA step from two 16-byte LDS operands
__device__ __forceinline__ f32x4 mfma32(s16x8 a, s16x8 b, f32x4 c) {
    return __builtin_amdgcn_mfma_f32_16x16x32_bf16(__builtin_bit_cast(bf16x8, a), __builtin_bit_cast(bf16x8, b), c, 0,
                                                   0, 0);
}

// round to nearest even
__device__ __forceinline__ short to_bf16(float x) {
    unsigned u = __float_as_uint(x);
    u += 0x7FFFu + ((u >> 16) & 1u);
    return (short)(u >> 16);
}

enum { FWD = 0, BWD_DATA = 1, BWD_W = 2 };

struct ConvArgs {
    int n, h, w, cin, cout;
    const float* x;      // [n][cin][h][w]
    const float* dy;     // [n][cout][h][w]
    const float* y;      // layer output for the ReLU mask
    const short* wb;     // [cout][cin*9] bf16
    const float* bias;   // [cout]
    float* out;
    int relu;
    int accumulate;
};

// LDS-tiled implicit GEMM for the forward and backward-data convolutions.
// A workgroup owns an 8x8 pixel tile of one image and up to 64 output
// channels; the K loop runs over 16-channel chunks, and inside a chunk over
// pairs of taps, so one MFMA k-step is (taps 2p and 2p+1, 16 channels): lane
// group g takes tap 2p + g/2 and channels 8*(g%2) .. +7 (the tenth tap of the
// last pair is zero).  Per chunk the input patch (10x10 pixels with the
// 1-pixel halo, channel fastest) and the weights ([64 rows][9 taps][16
// channels]) are converted to bf16 into LDS once; every lane then reads its 8
// k-values of A and of B as one 16-byte LDS word.  Waves take the (16-row,
// 16-pixel) output subtiles round-robin.
//   FWD      : in = x (cin channels), A[m][t][c] = W[m][c][t]
//   BWD_DATA : in = dy masked by y > 0 (cout channels), A[m][t][c] = W[c][m][8-t]
constexpr int TP = 8;             // pixel tile edge
constexpr int TPH = TP + 2;       // with halo
constexpr int CK = 16;            // channels per chunk
constexpr int MT = 64;            // output channels per workgroup

template <int MODE>
__global__ __launch_bounds__(256) void k_conv_tiled(ConvArgs a) {
    __shared__ __attribute__((aligned(16))) short patch[TPH * TPH * CK];
    __shared__ __attribute__((aligned(16))) short wl[MT * 9 * CK];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int Cin = MODE == FWD ? a.cin : a.cout;    // channels of the GEMM's input side
    const int M = MODE == FWD ? a.cout : a.cin;      // output channels
    const int tx = (a.w + TP - 1) / TP, ty = (a.h + TP - 1) / TP;
    const int tile = blockIdx.x, b = tile / (tx * ty), r = tile % (tx * ty);
    const int y0 = (r / tx) * TP, x0 = (r % tx) * TP;
    const int m0 = blockIdx.y * MT;
    const int Mb = min(4, (M - m0 + 15) / 16);       // 16-row blocks in this workgroup
    const long hw = (long)a.h * a.w;
    const float* in = MODE == FWD ? a.x : a.dy;
    const int li = lane & 15, lk = 4 * (lane >> 4), lg = lane >> 4, lc = 8 * (lg & 1);
    f32x4 acc[4];
#pragma unroll
    for (int q = 0; q < 4; ++q) acc[q] = f32x4{0.f, 0.f, 0.f, 0.f};
    for (int c0 = 0; c0 < Cin; c0 += CK) {
        __syncthreads();  // previous chunk's LDS reads are done
        for (int e = threadIdx.x; e < TPH * TPH * CK; e += 256) {
            const int px = e % TPH, py = (e / TPH) % TPH, c = e / (TPH * TPH);
            const int yy = y0 - 1 + py, xx = x0 - 1 + px, ch = c0 + c;
            float v = 0.f;
            if (ch < Cin && yy >= 0 && yy < a.h && xx >= 0 && xx < a.w) {
                const long gi = ((long)b * Cin + ch) * hw + (long)yy * a.w + xx;
                v = in[gi];
                if (MODE == BWD_DATA && a.relu && v != 0.f && a.y[gi] <= 0.f) v = 0.f;
            }
            patch[(py * TPH + px) * CK + c] = to_bf16(v);
        }
        for (int e = threadIdx.x; e < MT * 9 * CK; e += 256) {
            int m, c, t;
            short v = 0;
            if (MODE == FWD) {  // global [co][ci][t]: (m, c, t) with t fastest
                t = e % 9; c = (e / 9) % CK; m = e / (9 * CK);
                if (m0 + m < M && c0 + c < Cin) v = a.wb[((long)(m0 + m) * Cin + c0 + c) * 9 + t];
            } else {            // A[m][t][c] = W[c0+c][m0+m][8-t]; global [co][ci][t] with m = ci
                t = e % 9; m = (e / 9) % MT; c = e / (9 * MT);
                if (m0 + m < M && c0 + c < Cin) v = a.wb[((long)(c0 + c) * a.cin + m0 + m) * 9 + (8 - t)];
            }
            wl[(m * 9 + t) * CK + c] = v;
        }
        __syncthreads();
#pragma unroll
        for (int tp = 0; tp < 5; ++tp) {
            const int tr = 2 * tp + (lg >> 1), t = min(tr, 8);  // this lane group's tap (9: the zero tap)
            const int dy = t / 3, dx = t % 3;
            const s16x8 zero = s16x8{0, 0, 0, 0, 0, 0, 0, 0};
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                const int st = wave + 4 * q;  // subtile: 16-row block st/4, 16-pixel block st%4
                if (st >= Mb * 4) break;
                const int mb = st >> 2, nb = st & 3;
                const int n = nb * 16 + li, py = n / TP, px = n % TP;
                s16x8 av = *reinterpret_cast<const s16x8*>(&wl[((mb * 16 + li) * 9 + t) * CK + lc]);
                s16x8 bv = *reinterpret_cast<const s16x8*>(&patch[((py + dy) * TPH + px + dx) * CK + lc]);
                if (tr > 8) { av = zero; bv = zero; }
                acc[q] = mfma32(av, bv, acc[q]);
            }
        }
    }
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        const int st = wave + 4 * q;
        if (st >= Mb * 4) break;
        const int mb = st >> 2, nb = st & 3;
        const int n = nb * 16 + li, yy = y0 + n / TP, xx = x0 + n % TP;
        if (yy >= a.h || xx >= a.w) continue;
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const int m = m0 + mb * 16 + lk + j;
            if (m >= M) continue;
            const long o = ((long)b * M + m) * hw + (long)yy * a.w + xx;
            if (MODE == FWD) {
                float v = acc[q][j] + a.bias[m];
                a.out[o] = a.relu ? fmaxf(v, 0.f) : v;
            } else {
                a.out[o] = a.accumulate ? a.out[o] + acc[q][j] : acc[q][j];
            }
        }
    }
}

// Weight gradient dW[co][ci][t] = sum_p dY[co][p] X[ci][p + o(t)] as a GEMM
// with K = pixels: a workgroup owns up to 64 output channels x one 16-channel
// input chunk (144 (tap, channel) columns = 9 subtiles of 16) and strides over
// 8x8 pixel tiles (split-K over gridDim.z), accumulating in registers, with one
// atomic add per output at the end.  Per pixel tile, dY (ReLU-masked) goes to
// LDS as [co][64 pixels] and the input patch as three copies shifted by
// dx = 0, 1, 2 ([c][10 rows][8]).  One MFMA k-step is 32 pixels (4 rows of
// the tile): lane group g takes row 4s + g, so every lane reads its 8
// consecutive pixels of A and of B as one 16-byte LDS word.
__global__ __launch_bounds__(256) void k_conv_wgrad(ConvArgs a) {
    __shared__ __attribute__((aligned(16))) short dyl[MT * TP * TP];
    __shared__ __attribute__((aligned(16))) short xs[3 * CK * TPH * TP];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int li = lane & 15, lk = 4 * (lane >> 4), lg = lane >> 4;
    const int c0 = blockIdx.x * CK, m0 = blockIdx.y * MT;
    const int M = a.cout, Cin = a.cin;
    const int Mb = min(4, (M - m0 + 15) / 16);
    const int tx = (a.w + TP - 1) / TP, ty = (a.h + TP - 1) / TP, ntile = a.n * tx * ty;
    const long hw = (long)a.h * a.w;
    f32x4 acc[9];
#pragma unroll
    for (int q = 0; q < 9; ++q) acc[q] = f32x4{0.f, 0.f, 0.f, 0.f};
    for (int tile = blockIdx.z; tile < ntile; tile += gridDim.z) {
        const int b = tile / (tx * ty), r = tile % (tx * ty);
        const int y0 = (r / tx) * TP, x0 = (r % tx) * TP;
        __syncthreads();
        for (int e = threadIdx.x; e < MT * TP * TP; e += 256) {
            const int px = e % TP, py = (e / TP) % TP, m = e / (TP * TP);
            const int yy = y0 + py, xx = x0 + px;
            float v = 0.f;
            if (m0 + m < M && yy < a.h && xx < a.w) {
                const long gi = ((long)b * M + m0 + m) * hw + (long)yy * a.w + xx;
                v = a.dy[gi];
                if (a.relu && v != 0.f && a.y[gi] <= 0.f) v = 0.f;
            }
            dyl[m * TP * TP + py * TP + px] = to_bf16(v);
        }
        for (int e = threadIdx.x; e < CK * TPH * TPH; e += 256) {
            const int px = e % TPH, py = (e / TPH) % TPH, c = e / (TPH * TPH);
            const int yy = y0 - 1 + py, xx = x0 - 1 + px, ch = c0 + c;
            float v = 0.f;
            if (ch < Cin && yy >= 0 && yy < a.h && xx >= 0 && xx < a.w) v = a.x[((long)b * Cin + ch) * hw + (long)yy * a.w + xx];
            const short hv = to_bf16(v);
            // copy d holds column px - d of the tile's x range: xs[d][c][py][px - d]
#pragma unroll
            for (int d = 0; d < 3; ++d)
                if (px - d >= 0 && px - d < TP) xs[((d * CK + c) * TPH + py) * TP + px - d] = hv;
        }
        __syncthreads();
#pragma unroll
        for (int ks = 0; ks < 2; ++ks) {             // 32 pixels per k-step: rows 4ks .. 4ks+3
            const int py = ks * 4 + lg;              // this lane's row: 8 consecutive pixels
#pragma unroll
            for (int q = 0; q < 9; ++q) {
                const int st = wave + 4 * q;         // subtile = (16-row block, tap)
                if (st >= Mb * 9) break;
                const int mb = st / 9, t = st % 9, dy = t / 3, dx = t % 3;
                const s16x8 av = *reinterpret_cast<const s16x8*>(&dyl[(mb * 16 + li) * TP * TP + py * TP]);
                const s16x8 bv = *reinterpret_cast<const s16x8*>(&xs[((dx * CK + li) * TPH + py + dy) * TP]);
                acc[q] = mfma32(av, bv, acc[q]);
            }
        }
    }
#pragma unroll
    for (int q = 0; q < 9; ++q) {
        const int st = wave + 4 * q;
        if (st >= Mb * 9) break;
        const int mb = st / 9, t = st % 9;
        const int c = c0 + li;
        if (c >= Cin) continue;
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const int m = m0 + mb * 16 + lk + j;
            if (m < M) atomicAdd(&a.out[((long)m * Cin + c) * 9 + t], acc[q][j]);
        }
    }
}

// d bias[co] = sum over batch and pixels of the (ReLU-masked) output gradient;
// grid (channels, pixel chunks of 4096), one atomic add per block
__global__ void k_bias_grad(const float* dy, const float* y, int relu, int n, int c, int hw, float* db) {
    const int co = blockIdx.x;
    const long total = (long)n * hw;
    float s = 0.f;
    for (long e = (long)blockIdx.y * 4096 + threadIdx.x; e < min(total, (long)(blockIdx.y + 1) * 4096); e += blockDim.x) {
        const long b = e / hw, p = e % hw;
        const long i = (b * c + co) * hw + p;
        float g = dy[i];
        if (relu && y[i] <= 0.f) g = 0.f;
        s += g;
    }
    __shared__ float red[256];
    red[threadIdx.x] = s;
    __syncthreads();
    for (int o = blockDim.x / 2; o > 0; o >>= 1) {
        if ((int)threadIdx.x < o) red[threadIdx.x] += red[threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x == 0) atomicAdd(&db[co], red[0]);
}

// 2x2 average pooling; backward accumulates
__global__ void k_pool(const float* in, float* out, int nc, int h, int w) {
    const long i = blockIdx.x * (long)blockDim.x + threadIdx.x;
    const int ho = h / 2, wo = w / 2;
    if (i >= (long)nc * ho * wo) return;
    const int x = i % wo, y = (i / wo) % ho;
    const long c = i / ((long)ho * wo);
    const float* s = in + c * h * w + (2 * y) * w + 2 * x;
    out[i] = 0.25f * (s[0] + s[1] + s[w] + s[w + 1]);
}
__global__ void k_pool_bwd(const float* dout, float* din, int nc, int h, int w) {
    const long i = blockIdx.x * (long)blockDim.x + threadIdx.x;
    if (i >= (long)nc * h * w) return;
    const int x = i % w, y = (i / w) % h;
    const long c = i / ((long)h * w);
    din[i] += 0.25f * dout[c * (h / 2) * (w / 2) + (y / 2) * (w / 2) + x / 2];
}
// nearest 2x upsampling of [n][c][h][w] into channels [c0, c0+c) of a [n][ctot][2h][2w] buffer
__global__ void k_up(const float* in, float* out, int n, int c, int h, int w, int ctot, int c0) {
    const long i = blockIdx.x * (long)blockDim.x + threadIdx.x;
    const int H = 2 * h, W = 2 * w;
    if (i >= (long)n * c * H * W) return;
    const int x = i % W, y = (i / W) % H;
    const int ch = (i / ((long)H * W)) % c, b = i / ((long)c * H * W);
    out[(((long)b * ctot + c0 + ch) * H + y) * W + x] = in[(((long)b * c + ch) * h + y / 2) * w + x / 2];
}
__global__ void k_up_bwd(const float* dout, float* din, int n, int c, int h, int w, int ctot, int c0) {
    const long i = blockIdx.x * (long)blockDim.x + threadIdx.x;
    if (i >= (long)n * c * h * w) return;
    const int x = i % w, y = (i / w) % h;
    const int ch = (i / ((long)h * w)) % c, b = i / ((long)c * h * w);
    const int W = 2 * w;
    const float* s = dout + (((long)b * ctot + c0 + ch) * (2 * h) + 2 * y) * W + 2 * x;
    din[i] = s[0] + s[1] + s[W] + s[W + 1];
}
// copy channels: dst[:, d0:d0+c] (+)= src[:, s0:s0+c]
__global__ void k_chan_copy(const float* src, int sct, int s0, float* dst, int dct, int d0, int n, int c, int hw,
                            int add) {
    const long i = blockIdx.x * (long)blockDim.x + threadIdx.x;
    if (i >= (long)n * c * hw) return;
    const int p = i % hw, ch = (i / hw) % c, b = i / ((long)c * hw);
    const float v = src[((long)b * sct + s0 + ch) * hw + p];
    float* d = &dst[((long)b * dct + d0 + ch) * hw + p];
    *d = add ? *d + v : v;
}
__global__ void k_add(const float* a, const float* b, float* o, long n) {
    const long i = blockIdx.x * (long)blockDim.x + threadIdx.x;
    if (i < n) o[i] = a[i] + b[i];
}
// MSE: grad = 2 (y - t) / n, loss partial sums
__global__ void k_mse(const float* y, const float* t, float* g, long n, float* loss) {
    const long i = blockIdx.x * (long)blockDim.x + threadIdx.x;
    float d = 0.f;
    if (i < n) {
        d = y[i] - t[i];
        g[i] = 2.f * d / (float)n;
    }
    __shared__ float red[256];
    red[threadIdx.x] = d * d;
    __syncthreads();
    for (int o = blockDim.x / 2; o > 0; o >>= 1) {
        if ((int)threadIdx.x < o) red[threadIdx.x] += red[threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x == 0) atomicAdd(loss, red[0] / (float)n);
}
__global__ void k_adam(float* p, const float* g, float* m, float* v, short* pb, long n, float lr, float c1, float c2) {
    const long i = blockIdx.x * (long)blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float gi = g[i];
    const float mi = 0.9f * m[i] + 0.1f * gi, vi = 0.999f * v[i] + 0.001f * gi * gi;
    m[i] = mi;
    v[i] = vi;
    const float pi = p[i] - lr * (mi / c1) / (sqrtf(vi / c2) + 1e-8f);
    p[i] = pi;
    if (pb) pb[i] = to_bf16(pi);
}
__global__ void k_to_bf16(const float* p, short* pb, long n) {
    const long i = blockIdx.x * (long)blockDim.x + threadIdx.x;
    if (i < n) pb[i] = to_bf16(p[i]);
}
__global__ void k_zero(float* p, long n) {
    for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) p[i] = 0.f;
}

inline dim3 g1(long n) { return dim3((unsigned)((n + 255) / 256)); }

struct Conv {
    int cin, cout, relu;
    size_t woff, boff;  // offsets into the parameter vector
};

}  // namespace

struct erm_unet {
    int device = 0;
    hipStream_t stream = nullptr;
    int c[4];
    std::vector<Conv> conv;   // 16 layers in forward order
    size_t nparam = 0;
    float *p = nullptr, *g = nullptr, *m = nullptr, *v = nullptr;
    short* pb = nullptr;
    long adam_t = 0;
    // activation / gradient buffers for the current (n, h, w)
    int bn = 0, bh = 0, bw = 0;
    std::vector<float*> act, grad;
    float* xin = nullptr;
    float* tgt = nullptr;
    float* loss = nullptr;
    hipEvent_t ev[2] = {nullptr, nullptr};
    double last_ms = 0, last_flops = 0, flops_acc = 0;
    std::vector<size_t> act_size;
    int img_nx = 0, img_nz = 0;            // model image shape bound for erm_unet_precond
    std::vector<float> img_x, img_y;       // padded [1][2][H][W] staging
    std::vector<float> img_w;              // optional input weights [nx][nz] (erm_unet_bind_weights)

    ~erm_unet() {
        auto F = [](void* q) { if (q) (void)hipFree(q); };
        for (auto* a : act) F(a);
        for (auto* a : grad) F(a);
        F(p); F(g); F(m); F(v); F(pb); F(xin); F(tgt); F(loss);
        for (auto& e : ev) if (e) (void)hipEventDestroy(e);
        if (stream) (void)hipStreamDestroy(stream);
    }
};

namespace {

// buffer ids
enum {
    E1A, E1, P1, E2A, E2, P2, E3A, E3, P3, BA, BB, C3, D3A, D3, C2, D2A, D2, C1, D1A, D1, OUT, Y, NBUF
};

struct Shape { int c, lvl; };

void shapes(const erm_unet* u, Shape* s) {
    const int c1 = u->c[0], c2 = u->c[1], c3 = u->c[2], c4 = u->c[3];
    s[E1A] = {c1, 0}; s[E1] = {c1, 0}; s[P1] = {c1, 1};
    s[E2A] = {c2, 1}; s[E2] = {c2, 1}; s[P2] = {c2, 2};
    s[E3A] = {c3, 2}; s[E3] = {c3, 2}; s[P3] = {c3, 3};
    s[BA] = {c4, 3}; s[BB] = {c4, 3};
    s[C3] = {c4 + c3, 2}; s[D3A] = {c3, 2}; s[D3] = {c3, 2};
    s[C2] = {c3 + c2, 1}; s[D2A] = {c2, 1}; s[D2] = {c2, 1};
    s[C1] = {c2 + c1, 0}; s[D1A] = {c1, 0}; s[D1] = {c1, 0};
    s[OUT] = {2, 0}; s[Y] = {2, 0};
}

int ensure(erm_unet* u, int n, int h, int w) {
    if (h <= 0 || w <= 0 || h % 8 || w % 8 || n <= 0) {
        g_uerr = "h and w must be positive multiples of 8 and n > 0";
        return ERM_E_PARAM;
    }
    if (u->bn == n && u->bh == h && u->bw == w) return 0;
    for (auto* a : u->act) if (a) (void)hipFree(a);
    for (auto* a : u->grad) if (a) (void)hipFree(a);
    if (u->xin) (void)hipFree(u->xin);
    if (u->tgt) (void)hipFree(u->tgt);
    u->xin = u->tgt = nullptr;
    u->bn = u->bh = u->bw = 0;  // no valid buffers until every allocation below succeeds
    u->act.assign(NBUF, nullptr);
    u->grad.assign(NBUF, nullptr);
    u->act_size.assign(NBUF, 0);
    Shape s[NBUF];
    shapes(u, s);
    for (int b = 0; b < NBUF; ++b) {
        size_t sz = (size_t)n * s[b].c * (h >> s[b].lvl) * (w >> s[b].lvl);
        u->act_size[b] = sz;
        if (hipMalloc(&u->act[b], sz * 4) != hipSuccess || hipMalloc(&u->grad[b], sz * 4) != hipSuccess) {
            g_uerr = "hipMalloc failed";
            return ERM_E_NOMEM;
        }
    }
    size_t xs = (size_t)n * 2 * h * w * 4;
    if (hipMalloc(&u->xin, xs) != hipSuccess || hipMalloc(&u->tgt, xs) != hipSuccess) { g_uerr = "hipMalloc failed"; return ERM_E_NOMEM; }
    u->bn = n; u->bh = h; u->bw = w;
    return 0;
}

void conv_fwd(erm_unet* u, int li, const float* x, float* y, int n, int h, int w) {
    const Conv& L = u->conv[li];
    ConvArgs a{};
    a.n = n; a.h = h; a.w = w; a.cin = L.cin; a.cout = L.cout;
    a.x = x; a.wb = u->pb + L.woff; a.bias = u->p + L.boff; a.out = y; a.relu = L.relu;
    const long tiles = (long)n * ((h + TP - 1) / TP) * ((w + TP - 1) / TP);
    dim3 grd((unsigned)tiles, (unsigned)((L.cout + MT - 1) / MT));
    hipLaunchKernelGGL(k_conv_tiled<FWD>, grd, dim3(256), 0, u->stream, a);
    u->flops_acc += 2.0 * ((L.cout + 15) / 16 * 16) * (double)(tiles * TP * TP) * ((L.cin + CK - 1) / CK * CK * 9);
}

// backward of conv li: dy is the gradient w.r.t. its output y; adds dW/db into u->g,
// writes (or accumulates) dx when dx != nullptr
void conv_bwd(erm_unet* u, int li, const float* x, const float* y, const float* dy, float* dx, int acc_dx, int n,
              int h, int w) {
    const Conv& L = u->conv[li];
    ConvArgs a{};
    a.n = n; a.h = h; a.w = w; a.cin = L.cin; a.cout = L.cout;
    a.x = x; a.y = y; a.dy = dy; a.wb = u->pb + L.woff; a.relu = L.relu;
    const long P = (long)n * h * w;
    if (dx) {
        a.out = dx;
        a.accumulate = acc_dx;
        const long tiles = (long)n * ((h + TP - 1) / TP) * ((w + TP - 1) / TP);
        dim3 grd((unsigned)tiles, (unsigned)((L.cin + MT - 1) / MT));
        hipLaunchKernelGGL(k_conv_tiled<BWD_DATA>, grd, dim3(256), 0, u->stream, a);
        u->flops_acc += 2.0 * ((L.cin + 15) / 16 * 16) * (double)(tiles * TP * TP) * ((L.cout + CK - 1) / CK * CK * 9);
    }
    {
        a.out = u->g + L.woff;
        const int cchunks = (L.cin + CK - 1) / CK, mblocks = (L.cout + MT - 1) / MT;
        const long tiles = (long)n * ((h + TP - 1) / TP) * ((w + TP - 1) / TP);
        // split-K: about 1024 workgroups in all, at least 4 pixel tiles each
        const int ks = (int)std::max(1L, std::min(tiles / 4, 1024L / (cchunks * mblocks)));
        hipLaunchKernelGGL(k_conv_wgrad, dim3(cchunks, mblocks, ks), dim3(256), 0, u->stream, a);
        u->flops_acc += 2.0 * ((L.cout + 15) / 16 * 16) * (double)(cchunks * CK * 9) * (double)(tiles * TP * TP);
        const long chunks = (P + 4095) / 4096;
        hipLaunchKernelGGL(k_bias_grad, dim3(L.cout, (unsigned)chunks), dim3(256), 0, u->stream, dy, y, L.relu, n,
                           L.cout, h * w, u->g + L.boff);
    }
}

void forward(erm_unet* u, int n, int h, int w) {
    auto& A = u->act;
    const int c1 = u->c[0], c2 = u->c[1], c3 = u->c[2], c4 = u->c[3];
    int hh[4] = {h, h / 2, h / 4, h / 8}, ww[4] = {w, w / 2, w / 4, w / 8};
    conv_fwd(u, 0, u->xin, A[E1A], n, hh[0], ww[0]);
    conv_fwd(u, 1, A[E1A], A[E1], n, hh[0], ww[0]);
    hipLaunchKernelGGL(k_pool, g1((long)n * c1 * hh[1] * ww[1]), dim3(256), 0, u->stream, A[E1], A[P1], n * c1, hh[0], ww[0]);
    conv_fwd(u, 2, A[P1], A[E2A], n, hh[1], ww[1]);
    conv_fwd(u, 3, A[E2A], A[E2], n, hh[1], ww[1]);
    hipLaunchKernelGGL(k_pool, g1((long)n * c2 * hh[2] * ww[2]), dim3(256), 0, u->stream, A[E2], A[P2], n * c2, hh[1], ww[1]);
    conv_fwd(u, 4, A[P2], A[E3A], n, hh[2], ww[2]);
    conv_fwd(u, 5, A[E3A], A[E3], n, hh[2], ww[2]);
    hipLaunchKernelGGL(k_pool, g1((long)n * c3 * hh[3] * ww[3]), dim3(256), 0, u->stream, A[E3], A[P3], n * c3, hh[2], ww[2]);
    conv_fwd(u, 6, A[P3], A[BA], n, hh[3], ww[3]);
    conv_fwd(u, 7, A[BA], A[BB], n, hh[3], ww[3]);
    // decoder level 3
    hipLaunchKernelGGL(k_up, g1((long)n * c4 * hh[2] * ww[2]), dim3(256), 0, u->stream, A[BB], A[C3], n, c4, hh[3], ww[3], c4 + c3, 0);
    hipLaunchKernelGGL(k_chan_copy, g1((long)n * c3 * hh[2] * ww[2]), dim3(256), 0, u->stream, A[E3], c3, 0, A[C3], c4 + c3, c4, n, c3, hh[2] * ww[2], 0);
    conv_fwd(u, 8, A[C3], A[D3A], n, hh[2], ww[2]);
    conv_fwd(u, 9, A[D3A], A[D3], n, hh[2], ww[2]);
    hipLaunchKernelGGL(k_up, g1((long)n * c3 * hh[1] * ww[1]), dim3(256), 0, u->stream, A[D3], A[C2], n, c3, hh[2], ww[2], c3 + c2, 0);
    hipLaunchKernelGGL(k_chan_copy, g1((long)n * c2 * hh[1] * ww[1]), dim3(256), 0, u->stream, A[E2], c2, 0, A[C2], c3 + c2, c3, n, c2, hh[1] * ww[1], 0);
    conv_fwd(u, 10, A[C2], A[D2A], n, hh[1], ww[1]);
    conv_fwd(u, 11, A[D2A], A[D2], n, hh[1], ww[1]);
    hipLaunchKernelGGL(k_up, g1((long)n * c2 * hh[0] * ww[0]), dim3(256), 0, u->stream, A[D2], A[C1], n, c2, hh[1], ww[1], c2 + c1, 0);
    hipLaunchKernelGGL(k_chan_copy, g1((long)n * c1 * hh[0] * ww[0]), dim3(256), 0, u->stream, A[E1], c1, 0, A[C1], c2 + c1, c2, n, c1, hh[0] * ww[0], 0);
    conv_fwd(u, 12, A[C1], A[D1A], n, hh[0], ww[0]);
    conv_fwd(u, 13, A[D1A], A[D1], n, hh[0], ww[0]);
    conv_fwd(u, 14, A[D1], A[OUT], n, hh[0], ww[0]);
    hipLaunchKernelGGL(k_add, g1((long)n * 2 * h * w), dim3(256), 0, u->stream, A[OUT], u->xin, A[Y], (long)n * 2 * h * w);
}

// gradient w.r.t. Y must be in G[Y]; fills u->g
void backward(erm_unet* u, int n, int h, int w) {
    auto& A = u->act;
    auto& G = u->grad;
    const int c1 = u->c[0], c2 = u->c[1], c3 = u->c[2], c4 = u->c[3];
    int hh[4] = {h, h / 2, h / 4, h / 8}, ww[4] = {w, w / 2, w / 4, w / 8};
    hipLaunchKernelGGL(k_zero, dim3(256), dim3(256), 0, u->stream, u->g, (long)u->nparam);
    // Y = OUT + x: dOUT = dY
    conv_bwd(u, 14, A[D1], A[OUT], G[Y], G[D1], 0, n, hh[0], ww[0]);
    conv_bwd(u, 13, A[D1A], A[D1], G[D1], G[D1A], 0, n, hh[0], ww[0]);
    conv_bwd(u, 12, A[C1], A[D1A], G[D1A], G[C1], 0, n, hh[0], ww[0]);
    // C1 = [up(D2) | E1]
    hipLaunchKernelGGL(k_up_bwd, g1((long)n * c2 * hh[1] * ww[1]), dim3(256), 0, u->stream, G[C1], G[D2], n, c2, hh[1], ww[1], c2 + c1, 0);
    hipLaunchKernelGGL(k_chan_copy, g1((long)n * c1 * hh[0] * ww[0]), dim3(256), 0, u->stream, G[C1], c2 + c1, c2, G[E1], c1, 0, n, c1, hh[0] * ww[0], 0);
    conv_bwd(u, 11, A[D2A], A[D2], G[D2], G[D2A], 0, n, hh[1], ww[1]);
    conv_bwd(u, 10, A[C2], A[D2A], G[D2A], G[C2], 0, n, hh[1], ww[1]);
    hipLaunchKernelGGL(k_up_bwd, g1((long)n * c3 * hh[2] * ww[2]), dim3(256), 0, u->stream, G[C2], G[D3], n, c3, hh[2], ww[2], c3 + c2, 0);
    hipLaunchKernelGGL(k_chan_copy, g1((long)n * c2 * hh[1] * ww[1]), dim3(256), 0, u->stream, G[C2], c3 + c2, c3, G[E2], c2, 0, n, c2, hh[1] * ww[1], 0);
    conv_bwd(u, 9, A[D3A], A[D3], G[D3], G[D3A], 0, n, hh[2], ww[2]);
    conv_bwd(u, 8, A[C3], A[D3A], G[D3A], G[C3], 0, n, hh[2], ww[2]);
    hipLaunchKernelGGL(k_up_bwd, g1((long)n * c4 * hh[3] * ww[3]), dim3(256), 0, u->stream, G[C3], G[BB], n, c4, hh[3], ww[3], c4 + c3, 0);
    hipLaunchKernelGGL(k_chan_copy, g1((long)n * c3 * hh[2] * ww[2]), dim3(256), 0, u->stream, G[C3], c4 + c3, c4, G[E3], c3, 0, n, c3, hh[2] * ww[2], 0);
    conv_bwd(u, 7, A[BA], A[BB], G[BB], G[BA], 0, n, hh[3], ww[3]);
    conv_bwd(u, 6, A[P3], A[BA], G[BA], G[P3], 0, n, hh[3], ww[3]);
    // E3 feeds the skip (already in G[E3]) and the pool
    hipLaunchKernelGGL(k_pool_bwd, g1((long)n * c3 * hh[2] * ww[2]), dim3(256), 0, u->stream, G[P3], G[E3], n * c3, hh[2], ww[2]);
    conv_bwd(u, 5, A[E3A], A[E3], G[E3], G[E3A], 0, n, hh[2], ww[2]);
    conv_bwd(u, 4, A[P2], A[E3A], G[E3A], G[P2], 0, n, hh[2], ww[2]);
    hipLaunchKernelGGL(k_pool_bwd, g1((long)n * c2 * hh[1] * ww[1]), dim3(256), 0, u->stream, G[P2], G[E2], n * c2, hh[1], ww[1]);
    conv_bwd(u, 3, A[E2A], A[E2], G[E2], G[E2A], 0, n, hh[1], ww[1]);
    conv_bwd(u, 2, A[P1], A[E2A], G[E2A], G[P1], 0, n, hh[1], ww[1]);
    hipLaunchKernelGGL(k_pool_bwd, g1((long)n * c1 * hh[0] * ww[0]), dim3(256), 0, u->stream, G[P1], G[E1], n * c1, hh[0], ww[0]);
    conv_bwd(u, 1, A[E1A], A[E1], G[E1], G[E1A], 0, n, hh[0], ww[0]);
    conv_bwd(u, 0, u->xin, A[E1A], G[E1A], nullptr, 0, n, hh[0], ww[0]);
}

int upload(erm_unet* u, float* d, const float* h, size_t n) {
    if (hipMemcpyAsync(d, h, n * 4, hipMemcpyHostToDevice, u->stream) != hipSuccess) { g_uerr = "upload failed"; return ERM_E_DEVICE; }
    return 0;
}
int sync(erm_unet* u) {
    hipError_t e = hipStreamSynchronize(u->stream);
    if (e != hipSuccess) { g_uerr = std::string("device error: ") + hipGetErrorString(e); return ERM_E_DEVICE; }
    return 0;
}
int loss_and_grad(erm_unet* u, int n, int h, int w, float* loss_out) {
    const long ne = (long)n * 2 * h * w;
    hipLaunchKernelGGL(k_zero, dim3(1), dim3(64), 0, u->stream, u->loss, 1L);
    forward(u, n, h, w);
    hipLaunchKernelGGL(k_mse, g1(ne), dim3(256), 0, u->stream, u->act[Y], u->tgt, u->grad[Y], ne, u->loss);
    backward(u, n, h, w);
    if (loss_out) {
        if (hipMemcpyAsync(loss_out, u->loss, 4, hipMemcpyDeviceToHost, u->stream) != hipSuccess) { g_uerr = "copy failed"; return ERM_E_DEVICE; }
    }
    return sync(u);
}

}  // namespace

extern "C" {

const char* erm_unet_last_error(void) { return g_uerr.c_str(); }

int erm_unet_create(int device, int c1, int c2, int c3, int c4, unsigned long long seed, erm_unet** out) {
    if (!out || c1 <= 0 || c2 <= 0 || c3 <= 0 || c4 <= 0) { g_uerr = "bad arguments"; return ERM_E_PARAM; }
    *out = nullptr;
    if (hipSetDevice(device) != hipSuccess) { g_uerr = "hipSetDevice failed"; return ERM_E_DEVICE; }
    erm_unet* u = new erm_unet();
    u->device = device;
    u->c[0] = c1; u->c[1] = c2; u->c[2] = c3; u->c[3] = c4;
    const int spec[15][3] = {{2, c1, 1}, {c1, c1, 1}, {c1, c2, 1}, {c2, c2, 1}, {c2, c3, 1}, {c3, c3, 1},
                             {c3, c4, 1}, {c4, c4, 1}, {c4 + c3, c3, 1}, {c3, c3, 1}, {c3 + c2, c2, 1},
                             {c2, c2, 1}, {c2 + c1, c1, 1}, {c1, c1, 1}, {c1, 2, 0}};
    size_t off = 0;
    for (auto& s : spec) {
        Conv L;
        L.cin = s[0]; L.cout = s[1]; L.relu = s[2];
        L.woff = off;
        off += (size_t)L.cin * L.cout * 9;
        L.boff = off;
        off += L.cout;
        u->conv.push_back(L);
    }
    u->nparam = off;
    std::vector<float> init(off, 0.f);
    std::mt19937_64 rng(seed);
    for (size_t li = 0; li < u->conv.size(); ++li) {
        const Conv& L = u->conv[li];
        // He initialisation; the output layer starts at 1% scale so the net starts near identity
        const double sd = std::sqrt(2.0 / (L.cin * 9.0)) * (li + 1 == u->conv.size() ? 0.01 : 1.0);
        std::normal_distribution<double> N(0.0, sd);
        for (size_t i = 0; i < (size_t)L.cin * L.cout * 9; ++i) init[L.woff + i] = (float)N(rng);
    }
    bool ok = hipStreamCreateWithFlags(&u->stream, hipStreamNonBlocking) == hipSuccess &&
              hipMalloc(&u->p, off * 4) == hipSuccess && hipMalloc(&u->g, off * 4) == hipSuccess &&
              hipMalloc(&u->m, off * 4) == hipSuccess && hipMalloc(&u->v, off * 4) == hipSuccess &&
              hipMalloc(&u->pb, off * 2) == hipSuccess && hipMalloc(&u->loss, 4) == hipSuccess &&
              hipEventCreate(&u->ev[0]) == hipSuccess && hipEventCreate(&u->ev[1]) == hipSuccess;
    if (!ok) { g_uerr = "allocation failed"; delete u; return ERM_E_NOMEM; }
    int rc = erm_unet_set_params(u, init.data());
    if (rc) { delete u; return rc; }
    *out = u;
    return 0;
}

int erm_unet_destroy(erm_unet* u) {
    if (!u) return 0;
    (void)hipSetDevice(u->device);
    if (u->stream) (void)hipStreamSynchronize(u->stream);
    delete u;
    return 0;
}

size_t erm_unet_num_params(const erm_unet* u) { return u ? u->nparam : 0; }

int erm_unet_get_params(erm_unet* u, float* dst) {
    if (!u || !dst) return ERM_E_PARAM;
    if (hipMemcpyAsync(dst, u->p, u->nparam * 4, hipMemcpyDeviceToHost, u->stream) != hipSuccess) { g_uerr = "copy failed"; return ERM_E_DEVICE; }
    return sync(u);
}

int erm_unet_set_params(erm_unet* u, const float* src) {
    if (!u || !src) return ERM_E_PARAM;
    int rc = upload(u, u->p, src, u->nparam);
    if (rc) return rc;
    hipLaunchKernelGGL(k_to_bf16, g1((long)u->nparam), dim3(256), 0, u->stream, u->p, u->pb, (long)u->nparam);
    hipLaunchKernelGGL(k_zero, dim3(256), dim3(256), 0, u->stream, u->m, (long)u->nparam);
    hipLaunchKernelGGL(k_zero, dim3(256), dim3(256), 0, u->stream, u->v, (long)u->nparam);
    u->adam_t = 0;
    return sync(u);
}

int erm_unet_forward(erm_unet* u, const float* x, int n, int h, int w, float* y) {
    if (!u || !x || !y) return ERM_E_PARAM;
    (void)hipSetDevice(u->device);
    int rc = ensure(u, n, h, w);
    if (rc) return rc;
    if ((rc = upload(u, u->xin, x, (size_t)n * 2 * h * w))) return rc;
    u->flops_acc = 0;
    (void)hipEventRecord(u->ev[0], u->stream);
    forward(u, n, h, w);
    (void)hipEventRecord(u->ev[1], u->stream);
    if (hipMemcpyAsync(y, u->act[Y], (size_t)n * 2 * h * w * 4, hipMemcpyDeviceToHost, u->stream) != hipSuccess) { g_uerr = "copy failed"; return ERM_E_DEVICE; }
    if ((rc = sync(u))) return rc;
    float ms = 0;
    (void)hipEventElapsedTime(&ms, u->ev[0], u->ev[1]);
    u->last_ms = ms;
    u->last_flops = u->flops_acc;
    return 0;
}

int erm_unet_loss_grad(erm_unet* u, const float* x, const float* t, int n, int h, int w, float* loss, float* grad) {
    if (!u || !x || !t) return ERM_E_PARAM;
    (void)hipSetDevice(u->device);
    int rc = ensure(u, n, h, w);
    if (rc) return rc;
    const size_t ne = (size_t)n * 2 * h * w;
    if ((rc = upload(u, u->xin, x, ne)) || (rc = upload(u, u->tgt, t, ne))) return rc;
    if ((rc = loss_and_grad(u, n, h, w, loss))) return rc;
    if (grad) {
        if (hipMemcpyAsync(grad, u->g, u->nparam * 4, hipMemcpyDeviceToHost, u->stream) != hipSuccess) { g_uerr = "copy failed"; return ERM_E_DEVICE; }
        return sync(u);
    }
    return 0;
}

static int adam(erm_unet* u, float lr) {
    u->adam_t += 1;
    const float c1 = (float)(1.0 - std::pow(0.9, (double)u->adam_t)), c2 = (float)(1.0 - std::pow(0.999, (double)u->adam_t));
    hipLaunchKernelGGL(k_adam, g1((long)u->nparam), dim3(256), 0, u->stream, u->p, u->g, u->m, u->v, u->pb,
                       (long)u->nparam, lr, c1, c2);
    return 0;
}

int erm_unet_train_step(erm_unet* u, const float* x, const float* t, int n, int h, int w, float lr, float* loss) {
    if (!u || !x || !t) return ERM_E_PARAM;
    (void)hipSetDevice(u->device);
    int rc = ensure(u, n, h, w);
    if (rc) return rc;
    const size_t ne = (size_t)n * 2 * h * w;
    if ((rc = upload(u, u->xin, x, ne)) || (rc = upload(u, u->tgt, t, ne))) return rc;
    if ((rc = loss_and_grad(u, n, h, w, loss))) return rc;
    adam(u, lr);
    return sync(u);
}

int erm_unet_train(erm_unet* u, const float* x, const float* t, int count, int h, int w, int batch, int epochs,
                   float lr, unsigned long long seed, float* losses) {
    if (!u || !x || !t || count <= 0 || batch <= 0 || epochs <= 0) { g_uerr = "bad arguments"; return ERM_E_PARAM; }
    (void)hipSetDevice(u->device);
    const size_t per = (size_t)2 * h * w;
    std::vector<int> order(count);
    for (int i = 0; i < count; ++i) order[i] = i;
    std::mt19937_64 rng(seed);
    std::vector<float> bx(per * batch), bt(per * batch);
    u->flops_acc = 0;
    (void)hipEventRecord(u->ev[0], u->stream);
    for (int e = 0; e < epochs; ++e) {
        std::shuffle(order.begin(), order.end(), rng);
        double sum = 0;
        int nb = 0;
        for (int s = 0; s + batch <= count; s += batch) {
            for (int b = 0; b < batch; ++b) {
                std::memcpy(&bx[b * per], x + (size_t)order[s + b] * per, per * 4);
                std::memcpy(&bt[b * per], t + (size_t)order[s + b] * per, per * 4);
            }
            int rc = ensure(u, batch, h, w);
            if (rc) return rc;
            if ((rc = upload(u, u->xin, bx.data(), per * batch)) || (rc = upload(u, u->tgt, bt.data(), per * batch))) return rc;
            float l = 0;
            if ((rc = loss_and_grad(u, batch, h, w, &l))) return rc;
            adam(u, lr);
            sum += l;
            ++nb;
        }
        if (losses) losses[e] = nb ? (float)(sum / nb) : 0.f;
    }
    (void)hipEventRecord(u->ev[1], u->stream);
    int rc = sync(u);
    float ms = 0;
    (void)hipEventElapsedTime(&ms, u->ev[0], u->ev[1]);
    u->last_ms = ms;
    u->last_flops = u->flops_acc;
    return rc;
}

int erm_unet_bind_model(erm_unet* u, int nx, int nz) {
    if (!u || nx <= 0 || nz <= 0) { g_uerr = "bad arguments"; return ERM_E_PARAM; }
    u->img_nx = nx;
    u->img_nz = nz;
    const int H = (nx + 7) / 8 * 8, W = (nz + 7) / 8 * 8;
    u->img_x.assign((size_t)2 * H * W, 0.f);
    u->img_y.assign((size_t)2 * H * W, 0.f);
    u->img_w.clear();
    return 0;
}

int erm_unet_bind_weights(erm_unet* u, const float* w) {
    if (!u) { g_uerr = "bad arguments"; return ERM_E_PARAM; }
    if (!w) { u->img_w.clear(); return 0; }
    if (u->img_nx <= 0) { g_uerr = "erm_unet_bind_weights: bind a model shape first"; return ERM_E_STATE; }
    u->img_w.assign(w, w + (size_t)u->img_nx * u->img_nz);
    return 0;
}

int erm_unet_precond(const float* g, float* z, void* user) {
    erm_unet* u = (erm_unet*)user;
    if (!u || !g || !z || u->img_nx <= 0) { g_uerr = "erm_unet_precond: net not bound to a model shape"; return ERM_E_STATE; }
    const int nx = u->img_nx, nz = u->img_nz, H = (nx + 7) / 8 * 8, W = (nz + 7) / 8 * 8;
    const size_t mn = (size_t)nx * nz;
    // the net's input: w * g (w = 1 unless weights are bound)
    const bool wt = !u->img_w.empty();
    for (int ch = 0; ch < 3; ++ch)
        for (size_t i = 0; i < mn; ++i) z[ch * mn + i] = wt ? u->img_w[i] * g[ch * mn + i] : g[ch * mn + i];
    double ss = 0;
    for (size_t i = 0; i < 2 * mn; ++i) ss += (double)z[i] * z[i];
    const double rms = std::sqrt(ss / (2.0 * mn));
    if (rms == 0) return 0;
    std::fill(u->img_x.begin(), u->img_x.end(), 0.f);
    for (int ch = 0; ch < 2; ++ch)
        for (int i = 0; i < nx; ++i)
            for (int k = 0; k < nz; ++k)
                u->img_x[((size_t)ch * H + i) * W + k] = (float)(z[ch * mn + (size_t)i * nz + k] / rms);
    int rc = erm_unet_forward(u, u->img_x.data(), 1, H, W, u->img_y.data());
    if (rc) return rc;
    for (int ch = 0; ch < 2; ++ch)
        for (int i = 0; i < nx; ++i)
            for (int k = 0; k < nz; ++k)
                z[ch * mn + (size_t)i * nz + k] = (float)(u->img_y[((size_t)ch * H + i) * W + k] * rms);
    return 0;
}

int erm_unet_last_stats(const erm_unet* u, double* ms, double* flops) {
    if (!u) return ERM_E_PARAM;
    if (ms) *ms = u->last_ms;
    if (flops) *flops = u->last_flops;
    return 0;
}

}  // extern "C"
