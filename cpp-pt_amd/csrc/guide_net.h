// Body of the guide network for one hidden width, included by guide_mlp.h once
// per width inside its own namespace (no include guard on purpose).  Before each
// inclusion guide_mlp.h defines CPT_GUIDE_HID (hidden width: 64 or 128) and
// CPT_GUIDE_LOBES (von Mises-Fisher lobes: 8 or 16).  Width 64 keeps its BF16
// weights resident in LDS for inference; width 128 (186 KB of BF16 weights, more
// than a CU's 160 KB of LDS) runs inference in FP8 only (102 KB of e4m3 weights in
// LDS).  Training reads the BF16 weights through the caches at both widths.

constexpr int NIN = 7;          // pos.xyz, normal (theta, phi), wo (theta, phi), all in [0,1]
constexpr int NBINS = 32;
constexpr int K0 = NIN * NBINS; // 224 input features
constexpr int HID = CPT_GUIDE_HID;
constexpr int NLOBES = CPT_GUIDE_LOBES;
constexpr int NOUT = (5 * NLOBES + 15) / 16 * 16;  // 5 per lobe, padded to MFMA tiles: 48 or 80
constexpr int NTH = HID / 16;   // 16-wide column tiles of a hidden layer
constexpr int NTO = NOUT / 16;  // ... of the output layer
static_assert(HID % 16 == 0 && 4 * HID % 256 == 0, "hidden width: a multiple of 64");
constexpr int TILE = 64;        // samples per block tile
constexpr int PAD = 12;         // bf16 padding of LDS rows (12: fewest bank conflicts measured)
constexpr int LD0 = K0 + PAD;   // 236
constexpr int LDH = HID + PAD;  // 76
constexpr int LDO = NOUT + PAD; // 60 or 92

// parameter offsets in the flat FP32 vector: W_l (K_l x N_l, row-major), then b_l
constexpr int W0 = 0;
constexpr int B0 = W0 + K0 * HID;
constexpr int W1 = B0 + HID;
constexpr int B1 = W1 + HID * HID;
constexpr int W2 = B1 + HID;
constexpr int B2 = W2 + HID * HID;
constexpr int W3 = B2 + HID;
constexpr int B3 = W3 + HID * HID;
constexpr int W4 = B3 + HID;
constexpr int B4 = W4 + HID * NOUT;
constexpr int NPARAM = B4 + NOUT;  // 30000 or 88656

// transposed BF16 weights as stored in global memory and LDS: Wt_l[n][k]
constexpr int T0 = 0;
constexpr int T1 = T0 + HID * LD0;
constexpr int T2 = T1 + HID * LDH;
constexpr int T3 = T2 + HID * LDH;
constexpr int T4 = T3 + HID * LDH;
constexpr int NWT = T4 + NOUT * LDH;  // bf16 elements
// the per-wave inference's copy of the output layer, rows in out_orig order
constexpr int T4P = NWT;
constexpr int NWT_ALL = NWT + NOUT * LDH;  // bf16 elements of the device copy
// BF16 weights resident in LDS (width 64); otherwise read from global memory
constexpr bool WT_LDS = HID <= 64;
// The training kernel reads them through the caches at every width: its block
// also holds a 64-sample tile's features and four hidden layers, and without the
// weights (66 KB at width 64) two blocks fit a CU's LDS instead of one (width 64:
// 331 -> 214 us per step of 204,800 samples)
constexpr bool WT_LDS_TRAIN = false;

constexpr float KAPPA_LOG_MIN = -3.0f, KAPPA_LOG_MAX = 6.0f;
constexpr float TWO_PI = 6.28318530717958647692f;

typedef short s4 __attribute__((ext_vector_type(4)));
typedef float f4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ uint16_t f2bf(float f) {  // round to nearest even
    uint32_t u = __float_as_uint(f);
    u += 0x7fffu + ((u >> 16) & 1u);
    return (uint16_t)(u >> 16);
}
__device__ __forceinline__ float bf2f(uint16_t h) { return __uint_as_float((uint32_t)h << 16); }

// one-blob encoding: bin b of input x gets exp(-0.5 ((x*NBINS - (b+0.5)))^2)
__device__ __forceinline__ float blob(float x, int b) {
    float d = x * (float)NBINS - ((float)b + 0.5f);
    return __expf(-0.5f * d * d);
}

// N consecutive bins b .. b+N-1 of input x, times scale, with two exponentials (FP8
// inference):
// blob(x, b+k+1) = blob(x, b+k) exp(d_k - 1/2) for d_k = 32 x - (b + k + 1/2), and
// each ratio is the previous one times e^-1 (a few ulp from blob() per step).  A bin
// within 5.2 of the peak (the ones that are not 0 in FP8; farther ones are below
// 1.3e-6) follows at most 7 bins that are within 12.2 of it, so no value it depends
// on underflows; bins whose chain starts beyond the FP32 range come out 0.
template <int N>
__device__ __forceinline__ void blob_run(float x, int b, float scale, float* v) {
    constexpr float INV_E = 0.36787944117144233f;
    const float d0 = x * (float)NBINS - ((float)b + 0.5f);
    float val = scale * __expf(-0.5f * d0 * d0);
    float r = __expf(d0 - 0.5f);
#pragma unroll
    for (int k = 0; k < N; k++) {
        v[k] = val;
        val *= r;
        r *= INV_E;
    }
}

// acc[nt] += A(rows of this wave) x Wt^T, A row-major (lda), Wt n-major (ldb), on
// gfx950's v_mfma_f32_16x16x32_bf16: lane l supplies A[l%16][k0 + 8(l/16) + j] and
// B[k0 + 8(l/16) + j][n0 + l%16], j = 0..7 (layout checked on the device with a
// one-wave probe of exact small-integer products), each as two 8-byte halves (rows
// are 8-byte aligned).  KU: unroll of the k loop (1 when Wt is in global memory:
// fully unrolled, the hoisted global loads would not fit the register file)
typedef short s8v __attribute__((ext_vector_type(8)));
typedef __bf16 bf8v __attribute__((ext_vector_type(8)));
__device__ __forceinline__ bf8v ld_bf8(const uint16_t* p) {
    const s4 lo = *reinterpret_cast<const s4*>(p), hi = *reinterpret_cast<const s4*>(p + 4);
    const s8v v = __builtin_shufflevector(lo, hi, 0, 1, 2, 3, 4, 5, 6, 7);
    return __builtin_bit_cast(bf8v, v);
}
template <int K, int NT, int KU = K / 32>
__device__ __forceinline__ void gemm_rows_nk(const uint16_t* A, int lda, const uint16_t* Wt, int ldb, f4* acc,
                                             int lane) {
    static_assert(K % 32 == 0, "K in steps of 32");
#pragma unroll KU
    for (int kt = 0; kt < K / 32; kt++) {
        const int kk = kt * 32 + 8 * (lane >> 4);
        const bf8v a = ld_bf8(A + (lane & 15) * lda + kk);
#pragma unroll
        for (int nt = 0; nt < NT; nt++) {
            const bf8v b = ld_bf8(Wt + (nt * 16 + (lane & 15)) * ldb + kk);
            acc[nt] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, acc[nt], 0, 0, 0);
        }
    }
}

// acc[nt] += A x B with B stored k-major (Bk[k][n], ldb): 4 strided reads per lane
template <int K, int NT, int KU = K / 16>
__device__ __forceinline__ void gemm_rows_kn(const uint16_t* A, int lda, const uint16_t* Bk, int ldb, f4* acc,
                                             int lane) {
#pragma unroll KU
    for (int kt = 0; kt < K / 16; kt++) {
        const int kk = kt * 16 + 4 * (lane >> 4);
        s4 a = *reinterpret_cast<const s4*>(A + (lane & 15) * lda + kk);
#pragma unroll
        for (int nt = 0; nt < NT; nt++) {
            s4 b;
            const uint16_t* bp = Bk + kk * ldb + nt * 16 + (lane & 15);
            b[0] = (short)bp[0];
            b[1] = (short)bp[ldb];
            b[2] = (short)bp[2 * ldb];
            b[3] = (short)bp[3 * ldb];
            acc[nt] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, acc[nt], 0, 0, 0);
        }
    }
}

// Stored mixture of one sample, 3 values per lobe k (structure of arrays: value v
// of lobe k at dst[(v NLOBES + k) stride]):
//   v = 0: half2 (kappa mu_x, kappa mu_y)
//   v = 1: half2 (kappa mu_z, w~), w~ the lobe weight rounded to half
//   v = 2: c = log(w~ / S C(kappa')) - kappa' (float), S = sum of the w~, kappa' =
//          |kappa mu| from the stored halves, C(k) = k / (2 pi (1 - exp(-2 k)))
// so the mixture pdf is sum_k exp(kappa mu_k . w + c_k), and a sampler that picks
// lobe k with probability w~_k / S and draws a vMF direction of concentration
// kappa' about (kappa mu) / kappa' (all from the stored values) has exactly that
// density; rounding the network's mixture to halves changes the guide, not the
// agreement between sampling and pdf.
typedef _Float16 h2v __attribute__((ext_vector_type(2)));
__device__ __forceinline__ uint32_t pack_h2(float a, float b) {
    return __builtin_bit_cast(uint32_t, __builtin_amdgcn_cvt_pkrtz(a, b));
}
__device__ __forceinline__ float2 unpack_h2(uint32_t u) {
    const h2v h = __builtin_bit_cast(h2v, u);
    return make_float2((float)h.x, (float)h.y);
}
__device__ __forceinline__ float mix_log_term(float wh, float S, float kap) {
    return __logf(wh / S * kap / (TWO_PI * (1.0f - __expf(-2.0f * kap)))) - kap;
}
// the stored halves of one lobe: kappa mu, the weight w~ and kappa' = |kappa mu|
struct MixLobe {
    float kx, ky, kz, w, kap;
};
__device__ __forceinline__ MixLobe mix_lobe(uint32_t p0, uint32_t p1) {
    const float2 a = unpack_h2(p0), b = unpack_h2(p1);
    MixLobe m;
    m.kx = a.x;
    m.ky = a.y;
    m.kz = b.x;
    m.w = b.y;
    m.kap = sqrtf(m.kx * m.kx + m.ky * m.ky + m.kz * m.kz);
    return m;
}

struct Lobes {
    float pi[NLOBES], kappa[NLOBES], mx[NLOBES], my[NLOBES], mz[NLOBES];
};

// raw outputs -> mixture (softmax weights, kappa = exp(clamped), unit mean)
__device__ __forceinline__ void decode(const float* o, Lobes& L) {
    float mxv = -INFINITY;
#pragma unroll
    for (int k = 0; k < NLOBES; k++) mxv = fmaxf(mxv, o[5 * k]);
    float s = 0.0f;
#pragma unroll
    for (int k = 0; k < NLOBES; k++) {
        L.pi[k] = __expf(o[5 * k] - mxv);
        s += L.pi[k];
    }
    float inv = 1.0f / s;
#pragma unroll
    for (int k = 0; k < NLOBES; k++) {
        L.pi[k] *= inv;
        L.kappa[k] = __expf(fminf(fmaxf(o[5 * k + 1], KAPPA_LOG_MIN), KAPPA_LOG_MAX));
        float x = o[5 * k + 2], y = o[5 * k + 3], z = o[5 * k + 4];
        float n = sqrtf(x * x + y * y + z * z);
        if (n < 1e-12f) {
            x = 0.0f;
            y = 0.0f;
            z = 1.0f;
            n = 1.0f;
        }
        L.mx[k] = x / n;
        L.my[k] = y / n;
        L.mz[k] = z / n;
    }
}

// vMF density kappa / (2 pi (1 - e^{-2 kappa})) exp(kappa (mu.w - 1))
__device__ __forceinline__ float vmf_pdf(float kappa, float cosang) {
    return kappa / (TWO_PI * (1.0f - __expf(-2.0f * kappa))) * __expf(kappa * (cosang - 1.0f));
}

// network inputs in [0,1]: position in the scene box, normal and wo as (theta/pi, phi/2pi)
__device__ __forceinline__ void features(const float* bmin, const float* bext, float px, float py, float pz,
                                         float nx, float ny, float nz, float ox, float oy, float oz, float* in) {
    in[0] = fminf(fmaxf((px - bmin[0]) / bext[0], 0.0f), 1.0f);
    in[1] = fminf(fmaxf((py - bmin[1]) / bext[1], 0.0f), 1.0f);
    in[2] = fminf(fmaxf((pz - bmin[2]) / bext[2], 0.0f), 1.0f);
    in[3] = acosf(fminf(fmaxf(nz, -1.0f), 1.0f)) * (1.0f / 3.14159265f);
    in[4] = (atan2f(ny, nx) + 3.14159265f) * (1.0f / TWO_PI);
    in[5] = acosf(fminf(fmaxf(oz, -1.0f), 1.0f)) * (1.0f / 3.14159265f);
    in[6] = (atan2f(oy, ox) + 3.14159265f) * (1.0f / TWO_PI);
}

// The same inputs with cheaper angles, for per-bounce inference (the recorded
// training vertices keep features()): acos by Abramowitz and Stegun 4.4.45
// (reflected for negative arguments; 6.8e-5 rad at most, evaluated in float),
// atan on [0, 1] by Hastings' odd polynomial of degree 11, extended to atan2 by
// the octant symmetries (2e-6 rad at most).  Both differences stay below 3e-5 of
// an input's range, a thousandth of a one-blob bin.
__device__ __forceinline__ float acos_in(float x) {
    const float ax = fminf(fabsf(x), 1.0f);
    const float r = sqrtf(1.0f - ax) * (1.5707288f + ax * (-0.2121144f + ax * (0.0742610f + ax * -0.0187293f)));
    return x < 0.0f ? 3.14159265f - r : r;
}
__device__ __forceinline__ float atan2_in(float y, float x) {
    const float ax = fabsf(x), ay = fabsf(y);
    const float mx = fmaxf(ax, ay), mn = fminf(ax, ay);
    const float a = mx > 0.0f ? mn / mx : 0.0f;
    const float a2 = a * a;
    float r = a * (0.99997726f + a2 * (-0.33262347f + a2 * (0.19354346f + a2 * (-0.11643287f +
                                                                  a2 * (0.05265332f + a2 * -0.01172120f)))));
    if (ay > ax) r = 1.57079633f - r;
    if (x < 0.0f) r = 3.14159265f - r;
    return y < 0.0f ? -r : r;
}
__device__ __forceinline__ void features_fast(const float* bmin, const float* bext, float px, float py, float pz,
                                              float nx, float ny, float nz, float ox, float oy, float oz, float* in) {
    in[0] = fminf(fmaxf((px - bmin[0]) / bext[0], 0.0f), 1.0f);
    in[1] = fminf(fmaxf((py - bmin[1]) / bext[1], 0.0f), 1.0f);
    in[2] = fminf(fmaxf((pz - bmin[2]) / bext[2], 0.0f), 1.0f);
    in[3] = fminf(acos_in(nz) * (1.0f / 3.14159265f), 1.0f);
    in[4] = fminf(fmaxf((atan2_in(ny, nx) + 3.14159265f) * (1.0f / TWO_PI), 0.0f), 1.0f);
    in[5] = fminf(acos_in(oz) * (1.0f / 3.14159265f), 1.0f);
    in[6] = fminf(fmaxf((atan2_in(oy, ox) + 3.14159265f) * (1.0f / TWO_PI), 0.0f), 1.0f);
}

// --------------------------------------------------------------------------
// Shared-memory tile: weights (transposed bf16), biases, activations.
// WL: the weights are in the struct (else the kernels read them from global memory)
template <bool WL>
struct SmemFwdT {
    static constexpr bool kWL = WL;
    uint16_t wt[WL ? NWT : 8];
    float bias[4 * HID + NOUT];
    uint16_t x0[TILE * LD0];
    uint16_t h[4][TILE * LDH];
};
using SmemFwd = SmemFwdT<WT_LDS>;
using SmemTrain = SmemFwdT<WT_LDS_TRAIN>;

template <class SM>
__device__ __forceinline__ void load_weights(SM& sm, const uint16_t* gwt, const float* params) {
    const uint4* src = reinterpret_cast<const uint4*>(gwt);
    uint4* dst = reinterpret_cast<uint4*>(sm.wt);
    if constexpr (SM::kWL)
        for (int i = threadIdx.x; i < NWT / 8; i += blockDim.x) dst[i] = src[i];
    for (int i = threadIdx.x; i < 4 * HID + NOUT; i += blockDim.x) {
        int off = i < HID ? B0 + i : i < 2 * HID ? B1 + i - HID : i < 3 * HID ? B2 + i - 2 * HID
                : i < 4 * HID ? B3 + i - 3 * HID : B4 + i - 4 * HID;
        sm.bias[i] = params[off];
    }
}

// one-blob features of the tile's samples into sm.x0 (samples beyond n are zero)
// (4 consecutive bins of one input per thread step, one 8-byte store)
template <class SM>
__device__ __forceinline__ void encode_tile(SM& sm, const float* in, int n) {
    constexpr int CPS = K0 / 4;  // chunks per sample
    static_assert(NBINS % 4 == 0 && LD0 % 4 == 0, "8-byte chunks");
    for (int c = threadIdx.x; c < TILE * CPS; c += blockDim.x) {
        const int s = c / CPS, f = (c - s * CPS) * 4;
        const int d = f / NBINS, b = f - d * NBINS;
        uint2 w = {0u, 0u};
        if (s < n) {
            // each bin by its own exponential: training's BF16 features match the
            // oracle's rounding bin for bin (the ratio recurrence of blob_run moves a
            // few roundings, and the width-128 gradient is sensitive to that)
            const float x = in[s * NIN + d];
            w.x = (uint32_t)f2bf(blob(x, b)) | ((uint32_t)f2bf(blob(x, b + 1)) << 16);
            w.y = (uint32_t)f2bf(blob(x, b + 2)) | ((uint32_t)f2bf(blob(x, b + 3)) << 16);
        }
        *reinterpret_cast<uint2*>(sm.x0 + s * LD0 + f) = w;
    }
}

// First of the 16 bins an inference encode computes for input x: the bins
// c - 6 .. c + 5 around c = floor(32 x) (the peak lies in bin c) lie in
// [w0, w0 + 16), w0 a multiple of 4 in [0, 16].  Outside the window every bin is
// below e^-15.1 = 2.7e-7: 0 in FP8 (x 448 it rounds to 0), and in BF16 it moves a
// layer-0 pre-activation by less than 1e-6 of its size.
__device__ __forceinline__ int blob_window(float x) {
    const int c = (int)(x * (float)NBINS);
    return min(max(((c - 6) >> 2) << 2, 0), NBINS - 16);
}

// Inference encode (BF16): per (sample, input) pair the 32 bins are zeroed and the
// 16 of blob_window are written, each from its own exponential (samples beyond n
// are zero).  Training keeps encode_tile, every bin computed.
__device__ __forceinline__ void encode_tile_window(SmemFwd& sm, const float* in, int n) {
    static_assert(NBINS == 32 && LD0 % 4 == 0, "8-byte stores of 4 bins");
    for (int p = threadIdx.x; p < TILE * NIN; p += blockDim.x) {
        const int s = p / NIN, d = p - s * NIN;
        uint2* seg = reinterpret_cast<uint2*>(sm.x0 + s * LD0 + d * NBINS);
#pragma unroll
        for (int k = 0; k < NBINS / 4; k++) seg[k] = make_uint2(0u, 0u);
        if (s < n) {
            const float x = in[s * NIN + d];
            const int w0 = blob_window(x);
#pragma unroll
            for (int k = 0; k < 4; k++) {
                const int b = w0 + 4 * k;
                seg[b / 4] = make_uint2((uint32_t)f2bf(blob(x, b)) | ((uint32_t)f2bf(blob(x, b + 1)) << 16),
                                        (uint32_t)f2bf(blob(x, b + 2)) | ((uint32_t)f2bf(blob(x, b + 3)) << 16));
            }
        }
    }
}

// forward through all layers; out (TILE x NOUT fp32, row stride NOUT) receives raw
// outputs; wt = the transposed BF16 weights (sm.wt, or global memory when !SM::kWL)
template <class SM>
__device__ __forceinline__ void forward_tile(SM& sm, const uint16_t* wt, float* out) {
    constexpr bool WL = SM::kWL;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int r0 = 16 * wave;
    {
        f4 acc[NTH] = {};
        gemm_rows_nk<K0, NTH, WL ? K0 / 32 : 1>(sm.x0 + r0 * LD0, LD0, wt + T0, LD0, acc, lane);
#pragma unroll
        for (int nt = 0; nt < NTH; nt++)
#pragma unroll
            for (int j = 0; j < 4; j++) {
                int m = r0 + 4 * (lane >> 4) + j, n = nt * 16 + (lane & 15);
                float z = acc[nt][j] + sm.bias[n];
                sm.h[0][m * LDH + n] = f2bf(fmaxf(z, 0.0f));
            }
    }
    __syncthreads();
#pragma unroll
    for (int l = 1; l < 4; l++) {
        f4 acc[NTH] = {};
        const int toff = l == 1 ? T1 : l == 2 ? T2 : T3;
        gemm_rows_nk<HID, NTH, WL ? HID / 32 : 1>(sm.h[l - 1] + r0 * LDH, LDH, wt + toff, LDH, acc, lane);
#pragma unroll
        for (int nt = 0; nt < NTH; nt++)
#pragma unroll
            for (int j = 0; j < 4; j++) {
                int m = r0 + 4 * (lane >> 4) + j, n = nt * 16 + (lane & 15);
                float z = acc[nt][j] + sm.bias[l * HID + n];
                sm.h[l][m * LDH + n] = f2bf(fmaxf(z, 0.0f));
            }
        __syncthreads();
    }
    {
        f4 acc[NTO] = {};
        gemm_rows_nk<HID, NTO, WL ? HID / 32 : 1>(sm.h[3] + r0 * LDH, LDH, wt + T4, LDH, acc, lane);
#pragma unroll
        for (int nt = 0; nt < NTO; nt++)
#pragma unroll
            for (int j = 0; j < 4; j++) {
                int m = r0 + 4 * (lane >> 4) + j, n = nt * 16 + (lane & 15);
                out[m * NOUT + n] = acc[nt][j] + sm.bias[4 * HID + n];
            }
    }
    __syncthreads();
}

// d(-log q(w))/d outputs for one sample (scaled by wgt); returns -log q
__device__ __forceinline__ float loss_grad(const float* o, float wx, float wy, float wz, float wgt, float* g) {
    Lobes L;
    decode(o, L);
    float v[NLOBES], q = 0.0f;
#pragma unroll
    for (int k = 0; k < NLOBES; k++) {
        v[k] = vmf_pdf(L.kappa[k], L.mx[k] * wx + L.my[k] * wy + L.mz[k] * wz);
        q += L.pi[k] * v[k];
    }
    q = fmaxf(q, 1e-30f);
#pragma unroll
    for (int k = 0; k < NLOBES; k++) {
        float r = L.pi[k] * v[k] / q;  // responsibility
        g[5 * k] = -wgt * (r - L.pi[k]);
        float kap = L.kappa[k];
        float ls = o[5 * k + 1];
        float e2 = __expf(-2.0f * kap);
        float c = L.mx[k] * wx + L.my[k] * wy + L.mz[k] * wz;
        float dlogv_dk = 1.0f / kap - 2.0f * e2 / (1.0f - e2) + (c - 1.0f);
        bool inside = ls > KAPPA_LOG_MIN && ls < KAPPA_LOG_MAX;
        g[5 * k + 1] = inside ? -wgt * r * dlogv_dk * kap : 0.0f;
        float ux = o[5 * k + 2], uy = o[5 * k + 3], uz = o[5 * k + 4];
        float un = sqrtf(ux * ux + uy * uy + uz * uz);
        float f = un < 1e-12f ? 0.0f : -wgt * r * kap / un;
        g[5 * k + 2] = f * (wx - L.mx[k] * c);
        g[5 * k + 3] = f * (wy - L.my[k] * c);
        g[5 * k + 4] = f * (wz - L.mz[k] * c);
    }
#pragma unroll
    for (int k = 5 * NLOBES; k < NOUT; k++) g[k] = 0.0f;
    return -__logf(q);
}



// Training rows in global memory (bf16, one row per sample):
// [x0 224 | h0 64 | h1 64 | h2 64 | h3 64 | dz0 64 | dz1 64 | dz2 64 | dz3 64 | dz4 48]
constexpr int RX0 = 0, RH = K0, RDZ = K0 + 4 * HID, RDZ4 = RDZ + 4 * HID;
constexpr int ROW = RDZ4 + NOUT;  // 784 or 1328

// Training samples: 11 floats (7 inputs, direction xyz, weight = target / sampling pdf)
constexpr int SAMPLE_FLOATS = 11;

// Inference over n input vectors: raw outputs (n x NOUT).
__global__ __launch_bounds__(256) void k_mlp_forward(const uint16_t* gwt, const float* params, const float* in,
                                                     int n, float* out) {
    extern __shared__ __align__(16) unsigned char smem_raw[];
    SmemFwd& sm = *reinterpret_cast<SmemFwd*>(smem_raw);
    __shared__ __align__(16) float o_tile[TILE * NOUT];
    load_weights(sm, gwt, params);
    __syncthreads();
    for (int t = blockIdx.x; t * TILE < n; t += gridDim.x) {
        int nt = min(TILE, n - t * TILE);
        encode_tile_window(sm, in + (size_t)t * TILE * NIN, nt);
        __syncthreads();
        forward_tile(sm, WT_LDS ? sm.wt : gwt, o_tile);
        for (int i = threadIdx.x; i < nt * NOUT; i += blockDim.x) out[(size_t)t * TILE * NOUT + i] = o_tile[i];
        __syncthreads();
    }
}

// --------------------------------------------------------------------------
// FP8 inference (e4m3, OCP format of gfx950) on v_mfma_f32_16x16x32_fp8_fp8.
// Weights: per layer scaled by 448 / max|W_l| and rounded to e4m3 (k_quant8);
// one-blob features scaled by 448; each hidden layer's ReLU outputs scaled per
// 16-row group (one wave's rows of the tile) by the largest power of two that
// keeps their maximum <= 448.  FP32 accumulation; the accumulator is rescaled by
// the two inverse scales before the bias.  The conversions do not saturate (they
// return NaN above 448), so weights and features are clamped first and the
// activation scale never exceeds the range.  Lane layout of the 16x16x32 MFMA: lane l supplies
// A[l % 16][8 (l / 16) + j] and B[8 (l / 16) + j][l % 16], j = 0..7, as 8 bytes.
// Training keeps the BF16 path; only inference uses these weights.
// Byte padding of FP8 LDS rows.  Width 128: 16 keeps every row 16-byte aligned,
// so an MFMA operand is two 16-byte LDS reads whose offsets fit the instruction
// (no address arithmetic per tile) and the 16 rows of a read (68 dwords apart)
// fall in distinct banks: 4M-sample inference 1078 -> 1020 us.  Width 64 keeps
// 8 (its operands as four 8-byte reads; 16 costs it registers: 575 -> 581 us).
constexpr int PAD8 = HID > 64 ? 16 : 8;
constexpr int K0P = 256;            // layer-0 K padded to two MFMA steps of 128 (zero weights beyond K0)
static_assert(K0 <= K0P && K0P == 8 * NBINS, "layer-0 steps: 4 inputs of 32 bins each");
constexpr int LQ0 = K0P + PAD8;     // 264 or 272
constexpr int LQH = HID + PAD8;     // 72 or 144
constexpr int Q0 = 0;
constexpr int Q1 = Q0 + HID * LQ0;
constexpr int Q2 = Q1 + HID * LQH;
constexpr int Q3 = Q2 + HID * LQH;
constexpr int Q4 = Q3 + HID * LQH;
constexpr int NWQ = Q4 + NOUT * LQH;  // 34176 or 101632 bytes
static_assert(PAD8 % 16 != 0 || (LQ0 % 16 == 0 && LQH % 16 == 0 && NWQ % 16 == 0), "16-byte rows");
constexpr float FP8_MAX = 448.0f;

__device__ __forceinline__ uint32_t fp8x2(float a, float b) {
    a = fminf(fmaxf(a, -FP8_MAX), FP8_MAX);
    b = fminf(fmaxf(b, -FP8_MAX), FP8_MAX);
    return (uint32_t)__builtin_amdgcn_cvt_pk_fp8_f32(a, b, 0, false) & 0xffffu;
}
__device__ __forceinline__ uint8_t fp8x1(float a) { return (uint8_t)(fp8x2(a, 0.0f) & 0xffu); }

// FP8 inference runs one wave per tile of 16 samples: the block holds one copy
// of the FP8 weights in LDS and each wave its own tile state (inputs and one
// activation tile; the one-blob features are formed in registers), so after the
// weight load no wave waits for another.  Waves per block: 16 at width 64 (four
// per SIMD, 83 KB of LDS); 12 at width 128 (three per SIMD, 153 KB), whose
// forward pass and 16-lobe decode need more than the 128 registers of four.
#ifndef CPT_WAVES8_128
#define CPT_WAVES8_128 12
#endif
constexpr int WAVES8 = HID <= 64 ? 16 : CPT_WAVES8_128;
constexpr int LPG = NLOBES / 4;           // lobes per 16-lane group of the output tile
constexpr int PADG = 4 * NTO - 5 * LPG;   // output slots per group beyond its lobes
static_assert(NLOBES % 4 == 0 && PADG >= 0 && NOUT == 5 * NLOBES + 4 * PADG, "output tile layout");

// Output-layer row order of the FP8 copy.  Lane l of the output MFMA holds
// features 16 nt + 4 (l / 16) + j (nt < NTO, j < 4) of one sample; in this order the
// lanes of group g = l / 16 hold the 5 outputs of lobes g LPG .. g LPG + LPG - 1
// (then PADG unused outputs), so the mixture is decoded in registers.
// out_orig(f): network output stored at feature f; out_slot: its inverse.
__host__ __device__ constexpr int out_orig(int f) {
    const int g = (f % 16) / 4, q = 4 * (f / 16) + f % 4;
    return q < 5 * LPG ? 5 * (g * LPG + q / 5) + q % 5 : 5 * NLOBES + g * PADG + (q - 5 * LPG);
}
__host__ __device__ constexpr int out_slot(int n) {
    const int g = n < 5 * NLOBES ? n / 5 / LPG : (PADG > 0 ? (n - 5 * NLOBES) / PADG : 0);
    const int q = n < 5 * NLOBES ? 5 * (n / 5 % LPG) + n % 5 : 5 * LPG + (PADG > 0 ? (n - 5 * NLOBES) % PADG : 0);
    return 16 * (q / 4) + 4 * g + q % 4;
}

struct WaveState8 {
    float in[64 * NIN];   // inputs of up to 64 samples (one per lane when gathered)
    uint8_t h[16 * LQH];  // hidden activations; each layer overwrites its input
};
struct SmemFwd8 {
    uint8_t wt[NWQ];
    float bias[4 * HID + NOUT];  // output-layer biases in feature order (out_orig)
    float inv_w[8];              // 1 / weight scale per layer (5 used)
    WaveState8 w[WAVES8];
};
static_assert(sizeof(SmemFwd8) <= 160 * 1024, "FP8 inference LDS");
static_assert(PAD8 % 16 != 0 || (offsetof(SmemFwd8, w) % 16 == 0 && sizeof(WaveState8) % 16 == 0 &&
                                 offsetof(WaveState8, h) % 16 == 0),
              "16-byte aligned activation rows");

__device__ __forceinline__ void load_weights8(SmemFwd8& sm, const uint8_t* gwt8, const float* params,
                                              const float* inv_scale) {
    const uint4* src = reinterpret_cast<const uint4*>(gwt8);
    uint4* dst = reinterpret_cast<uint4*>(sm.wt);
    for (int i = threadIdx.x; i < NWQ / 16; i += blockDim.x) dst[i] = src[i];
    for (int i = threadIdx.x; i < 4 * HID + NOUT; i += blockDim.x) {
        int off = i < HID ? B0 + i : i < 2 * HID ? B1 + i - HID : i < 3 * HID ? B2 + i - 2 * HID
                : i < 4 * HID ? B3 + i - 3 * HID : B4 + out_orig(i - 4 * HID);
        sm.bias[i] = params[off];
    }
    if (threadIdx.x < 5) sm.inv_w[threadIdx.x] = inv_scale[threadIdx.x];
}

// the lanes of a wave see each other's LDS writes made before this point, and the
// compiler moves no LDS access across it
__device__ __forceinline__ void wave_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Layer-0 operand of one lane for the K step of 128 that covers inputs
// 4 step .. 4 step + 3: lane l supplies row l % 16, bytes 32 (l / 16) .. + 31 of
// the step, i.e. the 32 one-blob bins (x 448, e4m3) of input 4 step + l / 16 of
// sample row l % 16.  The 16 bins of blob_window come from two runs of 8
// (blob_run: two exponentials each; a longer run from a clamped window start
// could begin in FP32 underflow and carry the zero to the peak), every other bin
// is 0 (below e^-15.1, which e4m3 rounds to 0), and so are the padding input 7
// and rows at or beyond nrows.
typedef int i8v __attribute__((ext_vector_type(8)));
__device__ __forceinline__ i8v blob_operand8(const float* in16, int nrows, int step, int lane) {
    const int m = lane & 15, d = 4 * step + (lane >> 4);
    i8v a = {0, 0, 0, 0, 0, 0, 0, 0};
    if (m < nrows && d < NIN) {
        const float x = in16[m * NIN + d];
        const int w0 = blob_window(x);
        float v[16];
        blob_run<8>(x, w0, FP8_MAX, v);
        blob_run<8>(x, w0 + 8, FP8_MAX, v + 8);
        int wp[4];
#pragma unroll
        for (int k = 0; k < 4; k++) wp[k] = (int)(fp8x2(v[4 * k], v[4 * k + 1]) | (fp8x2(v[4 * k + 2], v[4 * k + 3]) << 16));
        const int q0 = w0 >> 2;  // first 4-bin word of the window (0..4)
#pragma unroll
        for (int i = 0; i < 8; i++) {
            const int k = i - q0;
            a[i] = k == 0 ? wp[0] : k == 1 ? wp[1] : k == 2 ? wp[2] : k == 3 ? wp[3] : 0;
        }
    }
    return a;
}

// acc[nt] += (A(rows of this wave) x Wt^T)^T: the weights are the MFMA's first
// operand, so lane l holds features nt 16 + 4 (l / 16) + j (j = 0..3) of sample row
// l % 16, four consecutive outputs of one sample for packed stores.  Steps of 128
// use gfx950's v_mfma_scale_f32_16x16x128_f8f6f4 (both operands e4m3, unit E8M0
// scales; lane l supplies row l % 16, bytes 32 (l / 16) .. 32 (l / 16) + 31 of the
// step, a layout checked on the device with a one-wave probe), the rest of K steps
// of 32 on v_mfma_f32_16x16x32_fp8_fp8 (8 bytes per lane at 8 (l / 16)).
__device__ __forceinline__ i8v ld32b(const uint8_t* p) {
    i8v v;
    if constexpr (PAD8 % 16 == 0) {  // rows 16-byte aligned
        const int4 q0 = *reinterpret_cast<const int4*>(p), q1 = *reinterpret_cast<const int4*>(p + 16);
        v[0] = q0.x; v[1] = q0.y; v[2] = q0.z; v[3] = q0.w; v[4] = q1.x; v[5] = q1.y; v[6] = q1.z; v[7] = q1.w;
    } else {
        const int2 q0 = *reinterpret_cast<const int2*>(p), q1 = *reinterpret_cast<const int2*>(p + 8);
        const int2 q2 = *reinterpret_cast<const int2*>(p + 16), q3 = *reinterpret_cast<const int2*>(p + 24);
        v[0] = q0.x; v[1] = q0.y; v[2] = q1.x; v[3] = q1.y; v[4] = q2.x; v[5] = q2.y; v[6] = q3.x; v[7] = q3.y;
    }
    return v;
}
template <int K, int NT>
__device__ __forceinline__ void gemm8_rows_nk_t(const uint8_t* A, int lda, const uint8_t* Wt, int ldb, f4* acc,
                                                int lane) {
    static_assert(K % 32 == 0, "K in steps of 32");
    constexpr int K128 = K / 128 * 128;
#pragma unroll
    for (int k0 = 0; k0 < K128; k0 += 128) {
        const int kk = k0 + 32 * (lane >> 4);
        const i8v a = ld32b(A + (lane & 15) * lda + kk);
#pragma unroll
        for (int nt = 0; nt < NT; nt++) {
            const i8v b = ld32b(Wt + (nt * 16 + (lane & 15)) * ldb + kk);
            acc[nt] = __builtin_amdgcn_mfma_scale_f32_16x16x128_f8f6f4(b, a, acc[nt], 0, 0, 0, 127, 0, 127);
        }
    }
#pragma unroll
    for (int k0 = K128; k0 < K; k0 += 32) {
        const int kk = k0 + 8 * (lane >> 4);
        const long a = *reinterpret_cast<const long*>(A + (lane & 15) * lda + kk);
#pragma unroll
        for (int nt = 0; nt < NT; nt++) {
            const long b = *reinterpret_cast<const long*>(Wt + (nt * 16 + (lane & 15)) * ldb + kk);
            acc[nt] = __builtin_amdgcn_mfma_f32_16x16x32_fp8_fp8(b, a, acc[nt], 0, 0, 0);
        }
    }
}

// Four values as packed e4m3 after division by rs (gfx950's
// v_cvt_scalef32_pk_fp8_f32: round to nearest even, no saturation, checked on the
// device with an in-bounds probe); the callers' rs keep every quotient <= 448.
typedef short fp8pk_t __attribute__((ext_vector_type(2)));
__device__ __forceinline__ uint32_t fp8x4_scaled(float v0, float v1, float v2, float v3, float rs) {
    fp8pk_t p = {0, 0};
    p = __builtin_amdgcn_cvt_scalef32_pk_fp8_f32(p, v0, v1, rs, false);
    p = __builtin_amdgcn_cvt_scalef32_pk_fp8_f32(p, v2, v3, rs, true);
    return __builtin_bit_cast(uint32_t, p);
}

// Exponent k of the power-of-two activation scale 2^k for a group maximum mx > 0:
// the largest k with mx 2^k <= 448 (mx = m 2^e, m in [0.5, 1): 448 / m lies in
// (448, 896], so k = 8 - e when m > 0.875, else 9 - e), capped at 126 so that 2^-k
// stays a normal float.  Scaling by a power of two is exact, so the scaled values
// need no clamp and the inverse scale is exact.
__device__ __forceinline__ int act_scale_exp(float mx) {
    const float m = __builtin_amdgcn_frexp_mantf(mx);
    const int e = __builtin_amdgcn_frexp_expf(mx);
    return min((m > 0.875f ? 8 : 9) - e, 126);
}

// Forward pass of the wave's 16 rows (inputs at in16, nrows valid) through all
// layers with FP8 operands.  o: raw outputs in the output MFMA's layout, lane l
// holding features 16 nt + 4 (l / 16) + j of sample row l % 16 (out_orig order).
__device__ __forceinline__ void wave_forward8(const SmemFwd8& sm, WaveState8& ws, const float* in16, int nrows,
                                              int lane, f4 (&o)[NTO]) {
    const int m = lane & 15;         // this lane's sample row
    const int n4 = 4 * (lane >> 4);  // its first feature within a 16-wide tile
    float inv_a = 1.0f / FP8_MAX;    // inverse scale of the current layer's input
#pragma unroll
    for (int l = 0; l < 4; l++) {
        f4 acc[NTH] = {};
        if (l == 0) {
#pragma unroll
            for (int step = 0; step < K0P / 128; step++) {
                const i8v a = blob_operand8(in16, nrows, step, lane);
                const int kk = 128 * step + 32 * (lane >> 4);
#pragma unroll
                for (int nt = 0; nt < NTH; nt++) {
                    const i8v b = ld32b(sm.wt + Q0 + (nt * 16 + m) * LQ0 + kk);
                    acc[nt] = __builtin_amdgcn_mfma_scale_f32_16x16x128_f8f6f4(b, a, acc[nt], 0, 0, 0, 127, 0, 127);
                }
            }
        } else {
            gemm8_rows_nk_t<HID, NTH>(ws.h, LQH, sm.wt + (l == 1 ? Q1 : l == 2 ? Q2 : Q3), LQH, acc, lane);
        }
        wave_sync();  // every lane has read its operand rows before any overwrites them
        const float sc = inv_a * sm.inv_w[l];
        float mx = 0.0f;
#pragma unroll
        for (int nt = 0; nt < NTH; nt++) {
            const float4 bv = *reinterpret_cast<const float4*>(sm.bias + l * HID + nt * 16 + n4);
            const float bb[4] = {bv.x, bv.y, bv.z, bv.w};
#pragma unroll
            for (int j = 0; j < 4; j++) {
                const float z = fmaxf(fmaf(acc[nt][j], sc, bb[j]), 0.0f);
                acc[nt][j] = z;
                mx = fmaxf(mx, z);
            }
        }
        // one power-of-two scale per wave (16 rows)
        for (int off = 32; off > 0; off >>= 1) mx = fmaxf(mx, __shfl_xor(mx, off));
        inv_a = mx > 0.0f ? __builtin_amdgcn_ldexpf(1.0f, -act_scale_exp(mx)) : 1.0f;
        uint8_t* hd = ws.h + m * LQH + n4;
#pragma unroll
        for (int nt = 0; nt < NTH; nt++)
            *reinterpret_cast<uint32_t*>(hd + nt * 16) = fp8x4_scaled(acc[nt][0], acc[nt][1], acc[nt][2], acc[nt][3], inv_a);
        wave_sync();  // the next layer reads what the other lanes just wrote
    }
#pragma unroll
    for (int nt = 0; nt < NTO; nt++) o[nt] = f4{0.0f, 0.0f, 0.0f, 0.0f};
    gemm8_rows_nk_t<HID, NTO>(ws.h, LQH, sm.wt + Q4, LQH, o, lane);
    const float sc = inv_a * sm.inv_w[4];
#pragma unroll
    for (int nt = 0; nt < NTO; nt++) {
        const float4 bv = *reinterpret_cast<const float4*>(sm.bias + 4 * HID + nt * 16 + n4);
        o[nt] = f4{fmaf(o[nt][0], sc, bv.x), fmaf(o[nt][1], sc, bv.y), fmaf(o[nt][2], sc, bv.z),
                   fmaf(o[nt][3], sc, bv.w)};
    }
    wave_sync();  // the wave's state may be overwritten from here on
}

// FP8 copy of the weights: one block per layer; inv_scale[l] = max|W_l| / 448
__global__ __launch_bounds__(256) void k_quant8(const float* params, uint8_t* gwt8, float* inv_scale) {
    __shared__ float red[256];
    const int l = blockIdx.x;
    const int off = l == 0 ? W0 : l == 1 ? W1 : l == 2 ? W2 : l == 3 ? W3 : W4;
    const int K = l == 0 ? K0 : HID, N = l == 4 ? NOUT : HID;
    const int q = l == 0 ? Q0 : l == 1 ? Q1 : l == 2 ? Q2 : l == 3 ? Q3 : Q4;
    const int ld = l == 0 ? LQ0 : LQH;
    float m = 0.0f;
    for (int i = threadIdx.x; i < K * N; i += blockDim.x) m = fmaxf(m, fabsf(params[off + i]));
    red[threadIdx.x] = m;
    __syncthreads();
    for (int st = 128; st > 0; st >>= 1) {
        if ((int)threadIdx.x < st) red[threadIdx.x] = fmaxf(red[threadIdx.x], red[threadIdx.x + st]);
        __syncthreads();
    }
    const float M = red[0];
    const float sc = M > 0.0f ? FP8_MAX / M : 1.0f;
    if (threadIdx.x == 0) inv_scale[l] = M > 0.0f ? M / FP8_MAX : 1.0f;
    for (int i = threadIdx.x; i < K * N; i += blockDim.x) {
        const int k = i / N, n = i - k * N;
        gwt8[q + (l == 4 ? out_slot(n) : n) * ld + k] = fp8x1(params[off + i] * sc);
    }
    for (int i = threadIdx.x; i < N * (ld - K); i += blockDim.x) {
        const int n = i / (ld - K), k = K + (i - n * (ld - K));
        gwt8[q + n * ld + k] = 0;
    }
}

// FP8 inference over n input vectors: raw outputs (n x NOUT, network order), one
// wave per tile of 16 consecutive inputs
__global__ __launch_bounds__(64 * WAVES8) void k_mlp_forward8(const uint8_t* gwt8, const float* params,
                                                              const float* inv_scale, const float* in, int n,
                                                              float* out) {
    extern __shared__ __align__(16) unsigned char smem_raw[];
    SmemFwd8& sm = *reinterpret_cast<SmemFwd8*>(smem_raw);
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    WaveState8& ws = sm.w[wave];
    load_weights8(sm, gwt8, params, inv_scale);
    __syncthreads();
    for (int t = blockIdx.x * WAVES8 + wave; t * 16 < n; t += gridDim.x * WAVES8) {
        const int nr = min(16, n - t * 16);
        for (int i = lane; i < 16 * NIN; i += 64) ws.in[i] = i < nr * NIN ? in[(size_t)t * 16 * NIN + i] : 0.0f;
        wave_sync();
        f4 o[NTO];
        wave_forward8(sm, ws, ws.in, nr, lane, o);
        const int m = lane & 15, n4 = 4 * (lane >> 4);
        if (m < nr)
#pragma unroll
            for (int nt = 0; nt < NTO; nt++)
#pragma unroll
                for (int j = 0; j < 4; j++) out[((size_t)t * 16 + m) * NOUT + out_orig(nt * 16 + n4 + j)] = o[nt][j];
    }
}

// Forward + loss gradient + backward-data for tiles of training samples; writes
// the per-sample rows used by the weight-gradient GEMMs and each block's loss sum.
// Per-sample rows for the weight gradients, stored feature-major: value f of
// sample s at rows[f * npad + s] (npad: n rounded up to TILE, padding samples
// zero), so a block's samples of one feature are contiguous.  Writes values f0 ..
// f0 + nf - 1 of the tile's samples from LDS rows src (ld elements apart), two
// samples per 4-byte store.
__device__ __forceinline__ void store_rows_t(uint16_t* rows, size_t npad, size_t s0, int nt, int f0, int nf,
                                             const uint16_t* src, int ld) {
    for (int i = threadIdx.x; i < nf * (TILE / 2); i += blockDim.x) {
        const int f = i / (TILE / 2), s = 2 * (i - f * (TILE / 2));
        const uint32_t lo = s < nt ? src[s * ld + f] : 0u, hi = s + 1 < nt ? src[(s + 1) * ld + f] : 0u;
        *reinterpret_cast<uint32_t*>(rows + (size_t)(f0 + f) * npad + s0 + s) = lo | (hi << 16);
    }
}

__global__ __launch_bounds__(256) void k_train_tile(const uint16_t* gwt, const float* params, const float* samples,
                                                    int n, int npad, const float* wscale, uint16_t* rows,
                                                    float* loss_part) {
    extern __shared__ __align__(16) unsigned char smem_raw[];
    SmemTrain& sm = *reinterpret_cast<SmemTrain*>(smem_raw);
    __shared__ float in_s[TILE * NIN];
    __shared__ float dir_s[TILE * 4];
    __shared__ float red[256];
    // outputs and their gradients: in the x0 rows once those are stored, if they fit
    constexpr int OT_BYTES = TILE * NOUT * 4 + TILE * LDO * 2;
    constexpr bool OT_IN_X0 = OT_BYTES <= TILE * LD0 * 2;
    __shared__ __align__(16) unsigned char ot_own[OT_IN_X0 ? 16 : OT_BYTES];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, r0 = 16 * wave;
    load_weights(sm, gwt, params);
    const uint16_t* wt = WT_LDS_TRAIN ? sm.wt : gwt;
    const float ws = *wscale;
    float loss_acc = 0.0f;
    __syncthreads();
    for (int t = blockIdx.x; t * TILE < n; t += gridDim.x) {
        const int nt = min(TILE, n - t * TILE);
        const size_t s0 = (size_t)t * TILE;
        for (int i = threadIdx.x; i < TILE; i += blockDim.x) {
            const float* sp = samples + (s0 + i) * SAMPLE_FLOATS;
            for (int d = 0; d < NIN; d++) in_s[i * NIN + d] = i < nt ? sp[d] : 0.0f;
            for (int d = 0; d < 4; d++) dir_s[i * 4 + d] = i < nt ? sp[NIN + d] : 0.0f;
        }
        __syncthreads();
        encode_tile(sm, in_s, nt);
        __syncthreads();
        store_rows_t(rows, npad, s0, nt, RX0, K0, sm.x0, LD0);
        __syncthreads();
        float* o_tile = reinterpret_cast<float*>(OT_IN_X0 ? reinterpret_cast<unsigned char*>(sm.x0) : ot_own);
        uint16_t* dz4 = reinterpret_cast<uint16_t*>(o_tile + TILE * NOUT);
        forward_tile(sm, wt, o_tile);
#pragma unroll
        for (int l = 0; l < 4; l++) store_rows_t(rows, npad, s0, nt, RH + l * HID, HID, sm.h[l], LDH);
        if (threadIdx.x < TILE) {
            int s = threadIdx.x;
            float g[NOUT];
            if (s < nt) {
                float l = loss_grad(o_tile + s * NOUT, dir_s[4 * s], dir_s[4 * s + 1], dir_s[4 * s + 2],
                                    dir_s[4 * s + 3] * ws, g);
                loss_acc += dir_s[4 * s + 3] * ws * l;
            } else {
                for (int k = 0; k < NOUT; k++) g[k] = 0.0f;
            }
            for (int k = 0; k < NOUT; k++) dz4[s * LDO + k] = f2bf(g[k]);
            for (int k = NOUT; k < LDO; k++) dz4[s * LDO + k] = 0;
        }
        __syncthreads();
        store_rows_t(rows, npad, s0, nt, RDZ4, NOUT, dz4, LDO);
        // backward-data: dz_{l-1} = (dz_l W_l^T) .* (h_{l-1} > 0), written over h_{l-1}
        {
            f4 acc[NTH] = {};
            gemm_rows_kn<NOUT, NTH, WT_LDS_TRAIN ? NOUT / 16 : 1>(dz4 + r0 * LDO, LDO, wt + T4, LDH, acc, lane);
#pragma unroll
            for (int ntl = 0; ntl < NTH; ntl++)
#pragma unroll
                for (int j = 0; j < 4; j++) {
                    int m = r0 + 4 * (lane >> 4) + j, u = ntl * 16 + (lane & 15);
                    float hv = bf2f(sm.h[3][m * LDH + u]);
                    acc[ntl][j] = hv > 0.0f ? acc[ntl][j] : 0.0f;
                }
            __syncthreads();  // all reads of h[3] as activations are done (rows already stored)
#pragma unroll
            for (int ntl = 0; ntl < NTH; ntl++)
#pragma unroll
                for (int j = 0; j < 4; j++) {
                    int m = r0 + 4 * (lane >> 4) + j, u = ntl * 16 + (lane & 15);
                    sm.h[3][m * LDH + u] = f2bf(acc[ntl][j]);
                }
            __syncthreads();
        }
#pragma unroll
        for (int l = 3; l >= 1; l--) {
            const int toff = l == 1 ? T1 : l == 2 ? T2 : T3;
            f4 acc[NTH] = {};
            gemm_rows_kn<HID, NTH, WT_LDS_TRAIN ? HID / 16 : 1>(sm.h[l] + r0 * LDH, LDH, wt + toff, LDH, acc, lane);
#pragma unroll
            for (int ntl = 0; ntl < NTH; ntl++)
#pragma unroll
                for (int j = 0; j < 4; j++) {
                    int m = r0 + 4 * (lane >> 4) + j, u = ntl * 16 + (lane & 15);
                    float hv = bf2f(sm.h[l - 1][m * LDH + u]);
                    acc[ntl][j] = hv > 0.0f ? acc[ntl][j] : 0.0f;
                }
            __syncthreads();
#pragma unroll
            for (int ntl = 0; ntl < NTH; ntl++)
#pragma unroll
                for (int j = 0; j < 4; j++) {
                    int m = r0 + 4 * (lane >> 4) + j, u = ntl * 16 + (lane & 15);
                    sm.h[l - 1][m * LDH + u] = f2bf(acc[ntl][j]);
                }
            __syncthreads();
        }
        // sm.h[l] now holds dz_l for l = 0..3
#pragma unroll
        for (int l = 0; l < 4; l++) store_rows_t(rows, npad, s0, nt, RDZ + l * HID, HID, sm.h[l], LDH);
        __syncthreads();
    }
    red[threadIdx.x] = loss_acc;
    __syncthreads();
    for (int off = 128; off > 0; off >>= 1) {
        if ((int)threadIdx.x < off) red[threadIdx.x] += red[threadIdx.x + off];
        __syncthreads();
    }
    if (threadIdx.x == 0) loss_part[blockIdx.x] = red[0];
}

// Weight gradients dW_l = sum_s a_l[s]^T dz_l[s] over a slice of the batch per
// block (MFMA on 16x16 output tiles spread over 4 waves; with DW_GROUPS > 1 the
// tiles are split further over blockIdx.y so each wave's accumulators stay in
// registers), bias gradients by plain sums (blockIdx.y == 0); each block writes
// its entries of its partial NPARAM vector.
constexpr int NTILES = (K0 / 16) * NTH + 3 * NTH * NTH + NTH * NTO;  // 116 or 344
constexpr int DW_GROUPS = HID <= 64 ? 4 : 8;
constexpr int TPW = (NTILES + 4 * DW_GROUPS - 1) / (4 * DW_GROUPS);  // 8 or 11 tiles per wave
constexpr int CH = 32;                 // samples per MFMA step of k_train_dw

__device__ __forceinline__ void tile_of(int t, int& layer, int& mt, int& nt) {
    constexpr int TL0 = (K0 / 16) * NTH, TLH = NTH * NTH;
    if (t < TL0) { layer = 0; mt = t / NTH; nt = t % NTH; return; }
    t -= TL0;
    if (t < 3 * TLH) { layer = 1 + t / TLH; mt = (t % TLH) / NTH; nt = t % NTH; return; }
    t -= 3 * TLH;
    layer = 4; mt = t / NTO; nt = t % NTO;
}

// grid (blocks, DW_GROUPS) x 256 threads.  rows is feature-major (store_rows_t);
// the block sums samples [blockIdx.x per_block, + per_block) of npad, per_block a
// multiple of 32.  Each 16x16 output tile takes one v_mfma_f32_16x16x32_bf16 per 32
// samples: lane l loads its 8 samples of activation row mt 16 + l % 16 and of
// gradient row nt 16 + l % 16 as one 16-byte load each (samples 8 (l / 16) ..).
__device__ __forceinline__ bf8v ld_bf8_16(const uint16_t* p) {
    return __builtin_bit_cast(bf8v, *reinterpret_cast<const uint4*>(p));
}
__global__ __launch_bounds__(256) void k_train_dw(const uint16_t* rows, int npad, int per_block, float* partial) {
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int tbase = blockIdx.y * 4 * TPW;
    const bool biases = blockIdx.y == 0;
    const int begin = min(npad, blockIdx.x * per_block);
    const int end = min(npad, begin + per_block);
    f4 acc[TPW];
#pragma unroll
    for (int i = 0; i < TPW; i++) acc[i] = f4{0, 0, 0, 0};
    for (int k0 = begin; k0 < end; k0 += 32) {
        const int ks = k0 + 8 * (lane >> 4);
#pragma unroll
        for (int i = 0; i < TPW; i++) {
            const int t = tbase + wave + 4 * i;
            if (t < NTILES) {
                int layer, mt, nt;
                tile_of(t, layer, mt, nt);
                const int aoff = layer == 0 ? RX0 : RH + (layer - 1) * HID;
                const int doff = layer == 4 ? RDZ4 : RDZ + layer * HID;
                const bf8v a = ld_bf8_16(rows + (size_t)(aoff + mt * 16 + (lane & 15)) * npad + ks);
                const bf8v b = ld_bf8_16(rows + (size_t)(doff + nt * 16 + (lane & 15)) * npad + ks);
                acc[i] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, acc[i], 0, 0, 0);
            }
        }
    }
    float* out = partial + (size_t)blockIdx.x * NPARAM;
#pragma unroll
    for (int i = 0; i < TPW; i++) {
        const int t = tbase + wave + 4 * i;
        if (t < NTILES) {
            int layer, mt, nt;
            tile_of(t, layer, mt, nt);
            const int N = layer == 4 ? NOUT : HID;
            const int woff = layer == 0 ? W0 : layer == 1 ? W1 : layer == 2 ? W2 : layer == 3 ? W3 : W4;
#pragma unroll
            for (int j = 0; j < 4; j++) {
                int k = mt * 16 + 4 * (lane >> 4) + j, nn = nt * 16 + (lane & 15);
                out[woff + k * N + nn] = acc[i][j];
            }
        }
    }
    if (biases) {
        // bias gradients: the block's samples of each gradient row, in sample order
        for (int e = threadIdx.x; e < 4 * HID + NOUT; e += blockDim.x) {
            const uint16_t* r = rows + (size_t)(RDZ + e) * npad;  // RDZ4 = RDZ + 4 HID
            float sum = 0.0f;
            for (int s = begin; s < end; s += 8) {
                const uint4 v = *reinterpret_cast<const uint4*>(r + s);
                const uint32_t w[4] = {v.x, v.y, v.z, v.w};
#pragma unroll
                for (int q = 0; q < 4; q++) {
                    sum += bf2f((uint16_t)(w[q] & 0xffffu));
                    sum += bf2f((uint16_t)(w[q] >> 16));
                }
            }
            const int l = e / HID, u = e % HID;
            out[e < 4 * HID ? (l == 0 ? B0 : l == 1 ? B1 : l == 2 ? B2 : B3) + u : B4 + (e - 4 * HID)] = sum;
        }
    }
}

// sum of the per-block partials in block order (deterministic)
__global__ void k_reduce_partials(const float* partial, int nblocks, int stride, int count, float* out) {
    for (int p = blockIdx.x * blockDim.x + threadIdx.x; p < count; p += gridDim.x * blockDim.x) {
        float s = 0.0f;
#pragma unroll 8
        for (int b = 0; b < nblocks; b++) s += partial[(size_t)b * stride + p];
        out[p] = s;
    }
}

// 1 / sum of sample weights, deterministic in two steps: per-block partial sums
// (grid-stride, fixed grid) and their sum in block order by one block.
__global__ __launch_bounds__(256) void k_weight_partial(const float* samples, int n, float* part) {
    __shared__ float red[256];
    float s = 0.0f;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
        s += samples[(size_t)i * SAMPLE_FLOATS + 10];
    red[threadIdx.x] = s;
    __syncthreads();
    for (int off = 128; off > 0; off >>= 1) {
        if ((int)threadIdx.x < off) red[threadIdx.x] += red[threadIdx.x + off];
        __syncthreads();
    }
    if (threadIdx.x == 0) part[blockIdx.x] = red[0];
}

__global__ __launch_bounds__(256) void k_weight_scale(const float* part, int nparts, float* wscale) {
    __shared__ float red[256];
    float s = 0.0f;
    for (int i = threadIdx.x; i < nparts; i += blockDim.x) s += part[i];
    red[threadIdx.x] = s;
    __syncthreads();
    for (int off = 128; off > 0; off >>= 1) {
        if ((int)threadIdx.x < off) red[threadIdx.x] += red[threadIdx.x + off];
        __syncthreads();
    }
    if (threadIdx.x == 0) *wscale = red[0] > 0.0f ? 1.0f / red[0] : 0.0f;
}

// Adam on the FP32 master weights, then refresh the transposed BF16 copy
__global__ void k_adam(float* params, float* m, float* v, const float* grad, int step, float lr, float b1, float b2,
                       float eps, uint16_t* gwt) {
    const float c1 = 1.0f / (1.0f - powf(b1, (float)step));
    const float c2 = 1.0f / (1.0f - powf(b2, (float)step));
    for (int p = blockIdx.x * blockDim.x + threadIdx.x; p < NPARAM; p += gridDim.x * blockDim.x) {
        float g = grad[p];
        if (!isfinite(g)) g = 0.0f;
        float mm = b1 * m[p] + (1.0f - b1) * g;
        float vv = b2 * v[p] + (1.0f - b2) * g * g;
        m[p] = mm;
        v[p] = vv;
        float w = params[p] - lr * (mm * c1) / (sqrtf(vv * c2) + eps);
        params[p] = w;
        int off, K, N, ld;
        if (p < B0) { off = W0; K = K0; N = HID; ld = LD0; (void)K; gwt[T0 + ((p - off) % N) * ld + (p - off) / N] = f2bf(w); }
        else if (p >= W1 && p < B1) { off = W1; N = HID; ld = LDH; gwt[T1 + ((p - off) % N) * ld + (p - off) / N] = f2bf(w); }
        else if (p >= W2 && p < B2) { off = W2; N = HID; ld = LDH; gwt[T2 + ((p - off) % N) * ld + (p - off) / N] = f2bf(w); }
        else if (p >= W3 && p < B3) { off = W3; N = HID; ld = LDH; gwt[T3 + ((p - off) % N) * ld + (p - off) / N] = f2bf(w); }
        else if (p >= W4 && p < B4) {
            off = W4; N = NOUT; ld = LDH;
            const uint16_t wb = f2bf(w);
            gwt[T4 + ((p - off) % N) * ld + (p - off) / N] = wb;
            gwt[T4P + out_slot((p - off) % N) * ld + (p - off) / N] = wb;  // inference copy
        }
    }
}

// BF16 transposed copy of all weights (after set_weights)
__global__ void k_pack_weights(const float* params, uint16_t* gwt) {
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < NWT_ALL; i += gridDim.x * blockDim.x) gwt[i] = 0;
    __syncthreads();
}
__global__ void k_pack_weights2(const float* params, uint16_t* gwt) {
    for (int p = blockIdx.x * blockDim.x + threadIdx.x; p < NPARAM; p += gridDim.x * blockDim.x) {
        float w = params[p];
        if (p < B0) gwt[T0 + ((p - W0) % HID) * LD0 + (p - W0) / HID] = f2bf(w);
        else if (p >= W1 && p < B1) gwt[T1 + ((p - W1) % HID) * LDH + (p - W1) / HID] = f2bf(w);
        else if (p >= W2 && p < B2) gwt[T2 + ((p - W2) % HID) * LDH + (p - W2) / HID] = f2bf(w);
        else if (p >= W3 && p < B3) gwt[T3 + ((p - W3) % HID) * LDH + (p - W3) / HID] = f2bf(w);
        else if (p >= W4 && p < B4) {
            gwt[T4 + ((p - W4) % NOUT) * LDH + (p - W4) / NOUT] = f2bf(w);
            gwt[T4P + out_slot((p - W4) % NOUT) * LDH + (p - W4) / NOUT] = f2bf(w);
        }
    }
}

// Mixtures of a tile's samples from their raw outputs, as decode() but spread over
// 4 consecutive threads per sample (NLOBES / 4 lobes each; the softmax maximum and
// sum, and the sum S of the stored weights, are combined across the 4 lanes); needs
// blockDim.x >= 4 TILE.  Stored as the 3 values per lobe described at MixLobe, value
// v of sample s at dst[v * stride + s] (structure of arrays).
__device__ __forceinline__ void decode_tile(const float* o_tile, int nt, float* dst, size_t stride) {
    constexpr int LPT = NLOBES / 4;
    static_assert(NLOBES % 4 == 0, "lobes per thread");
    const int s = threadIdx.x >> 2, k0 = (threadIdx.x & 3) * LPT;
    if (s >= nt) return;  // the 4 lanes of a sample share s
    const float* o = o_tile + s * NOUT;
    float mxv = -INFINITY;
#pragma unroll
    for (int j = 0; j < LPT; j++) mxv = fmaxf(mxv, o[5 * (k0 + j)]);
    mxv = fmaxf(mxv, __shfl_xor(mxv, 1));
    mxv = fmaxf(mxv, __shfl_xor(mxv, 2));
    float e[LPT], sum = 0.0f;
#pragma unroll
    for (int j = 0; j < LPT; j++) {
        e[j] = __expf(o[5 * (k0 + j)] - mxv);
        sum += e[j];
    }
    sum += __shfl_xor(sum, 1);
    sum += __shfl_xor(sum, 2);
    const float inv = 1.0f / sum;
    uint32_t p0[LPT], p1[LPT];
    float wsum = 0.0f;
#pragma unroll
    for (int j = 0; j < LPT; j++) {
        const int k = k0 + j;
        float x = o[5 * k + 2], y = o[5 * k + 3], z = o[5 * k + 4];
        float nrm = sqrtf(x * x + y * y + z * z);
        if (nrm < 1e-12f) {
            x = 0.0f;
            y = 0.0f;
            z = 1.0f;
            nrm = 1.0f;
        }
        const float kappa = __expf(fminf(fmaxf(o[5 * k + 1], KAPPA_LOG_MIN), KAPPA_LOG_MAX));
        const float kn = kappa / nrm;
        p0[j] = pack_h2(kn * x, kn * y);
        p1[j] = pack_h2(kn * z, e[j] * inv);
        wsum += unpack_h2(p1[j]).y;
    }
    wsum += __shfl_xor(wsum, 1);
    wsum += __shfl_xor(wsum, 2);
    float* d = dst + s;
#pragma unroll
    for (int j = 0; j < LPT; j++) {
        const int k = k0 + j;
        const MixLobe m = mix_lobe(p0[j], p1[j]);
        d[k * stride] = __uint_as_float(p0[j]);
        d[(NLOBES + k) * stride] = __uint_as_float(p1[j]);
        d[(2 * NLOBES + k) * stride] = mix_log_term(m.w, wsum, m.kap);
    }
}

// The path tracer's per-bounce inference, BF16 block tiles (width 128): weights
// into LDS once per block (or read through the L2), then per tile of up to TILE
// samples (NIN features each at `in`) the forward pass and the stored mixtures,
// sample s's lobe k at dst[{k, NLOBES + k, 2 NLOBES + k} stride + s].
__device__ __forceinline__ void infer_tile(SmemFwd& sm, const uint16_t* gwt, float* o_tile, const float* in, int nt,
                                           float* dst, size_t stride) {
    encode_tile_window(sm, in, nt);
    __syncthreads();
    forward_tile(sm, WT_LDS ? sm.wt : gwt, o_tile);
    decode_tile(o_tile, nt, dst, stride);
    __syncthreads();
}

// The mixtures of a wave's 16 rows from the registers of the per-wave forward
// passes, with the same arithmetic as decode_tile: the lanes of group g = lane / 16
// hold lobes g LPG .. g LPG + LPG - 1 of sample row lane % 16, and the softmax
// maximum and sum and the weight sum S are combined across the 4 groups (the same
// pairs as decode_tile's 4 threads).
__device__ __forceinline__ void wave_decode(const f4 (&o)[NTO], int lane, int nrows, float* dst, size_t stride) {
    const int g = lane >> 4, row = lane & 15;
    float v[4 * NTO];
#pragma unroll
    for (int nt = 0; nt < NTO; nt++)
#pragma unroll
        for (int j = 0; j < 4; j++) v[4 * nt + j] = o[nt][j];
    float mxv = -INFINITY;
#pragma unroll
    for (int i = 0; i < LPG; i++) mxv = fmaxf(mxv, v[5 * i]);
    mxv = fmaxf(mxv, __shfl_xor(mxv, 16));
    mxv = fmaxf(mxv, __shfl_xor(mxv, 32));
    float e[LPG], sum = 0.0f;
#pragma unroll
    for (int i = 0; i < LPG; i++) {
        e[i] = __expf(v[5 * i] - mxv);
        sum += e[i];
    }
    sum += __shfl_xor(sum, 16);
    sum += __shfl_xor(sum, 32);
    const float inv = 1.0f / sum;
    uint32_t p0[LPG], p1[LPG];
    float wsum = 0.0f;
#pragma unroll
    for (int i = 0; i < LPG; i++) {
        float x = v[5 * i + 2], y = v[5 * i + 3], z = v[5 * i + 4];
        float nrm = sqrtf(x * x + y * y + z * z);
        if (nrm < 1e-12f) {
            x = 0.0f;
            y = 0.0f;
            z = 1.0f;
            nrm = 1.0f;
        }
        const float kappa = __expf(fminf(fmaxf(v[5 * i + 1], KAPPA_LOG_MIN), KAPPA_LOG_MAX));
        const float kn = kappa / nrm;
        p0[i] = pack_h2(kn * x, kn * y);
        p1[i] = pack_h2(kn * z, e[i] * inv);
        wsum += unpack_h2(p1[i]).y;
    }
    wsum += __shfl_xor(wsum, 16);
    wsum += __shfl_xor(wsum, 32);
    if (row >= nrows) return;
    float* d = dst + row;
#pragma unroll
    for (int i = 0; i < LPG; i++) {
        const int k = g * LPG + i;
        const MixLobe m = mix_lobe(p0[i], p1[i]);
        d[k * stride] = __uint_as_float(p0[i]);
        d[(NLOBES + k) * stride] = __uint_as_float(p1[i]);
        d[(2 * NLOBES + k) * stride] = mix_log_term(m.w, wsum, m.kap);
    }
}

// FP8 mixtures of up to 64 samples whose inputs are in ws.in (nin valid), by one
// wave, four tiles of 16; sample s at dst + s as decode_tile
__device__ __forceinline__ void wave_infer8(const SmemFwd8& sm, WaveState8& ws, int nin, int lane, float* dst,
                                            size_t stride) {
    for (int q = 0; q < 4 && 16 * q < nin; q++) {
        const int nr = min(16, nin - 16 * q);
        f4 o[NTO];
        wave_forward8(sm, ws, ws.in + 16 * q * NIN, nr, lane, o);
        wave_decode(o, lane, nr, dst + 16 * q, stride);
    }
}

// FP8 mixtures of n input vectors (n x NIN) into dst (value v of sample s at
// dst[v * stride + s]), the code the path tracer runs per bounce after its own
// gather of the inputs; for tests of the mixtures it produces
__global__ __launch_bounds__(64 * WAVES8) void k_mlp_mixtures8(const uint8_t* gwt8, const float* params,
                                                               const float* inv_scale, const float* in, int n,
                                                               float* dst, size_t stride) {
    extern __shared__ __align__(16) unsigned char smem_raw[];
    SmemFwd8& sm = *reinterpret_cast<SmemFwd8*>(smem_raw);
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    WaveState8& ws = sm.w[wave];
    load_weights8(sm, gwt8, params, inv_scale);
    __syncthreads();
    for (int c = blockIdx.x * WAVES8 + wave; c * 64 < n; c += gridDim.x * WAVES8) {
        const int nin = min(64, n - c * 64);
        for (int i = lane; i < 64 * NIN; i += 64) ws.in[i] = i < nin * NIN ? in[(size_t)c * 64 * NIN + i] : 0.0f;
        wave_sync();
        wave_infer8(sm, ws, nin, lane, dst + (size_t)c * 64, stride);
    }
}

// --------------------------------------------------------------------------
// BF16 inference, one wave per 16-sample tile, at the width whose weights fit in
// LDS (64): the block holds one copy of the transposed BF16 weights, the output
// layer's rows in out_orig order (from the copy at T4P), and each wave its own tile
// state.  The one-blob features are formed in registers as the MFMA's second
// operand (the 16 bins of blob_window by the ratio recurrence, within a few ulp of
// blob() before their BF16 rounding; 0 elsewhere), the weights are the first
// operand, so a lane holds four consecutive outputs of one sample and the mixture is
// decoded in registers (wave_decode).  Per output the arithmetic of forward_tile
// (acc + bias, BF16 activations); only the MFMA operand order differs.
constexpr int WAVES16 = 12;
struct WaveState16 {
    float in[64 * NIN];    // inputs of up to 64 samples
    uint16_t h[16 * LDH];  // hidden activations; each layer overwrites its input
};
struct SmemInf16 {
    uint16_t wt[NWT];            // layers 0-3 as T0..T3, layer 4 from T4P
    float bias[4 * HID + NOUT];  // output-layer biases in out_orig order
    WaveState16 w[WAVES16];
};
static_assert(!WT_LDS || sizeof(SmemInf16) <= 160 * 1024, "BF16 inference LDS");

__device__ __forceinline__ void load_weights16(SmemInf16& sm, const uint16_t* gwt, const float* params) {
    static_assert(T4 % 8 == 0 && NWT % 8 == 0 && T4P % 8 == 0, "16-byte copies");
    const uint4* src = reinterpret_cast<const uint4*>(gwt);
    uint4* dst = reinterpret_cast<uint4*>(sm.wt);
    for (int i = threadIdx.x; i < NWT / 8; i += blockDim.x) dst[i] = src[i < T4 / 8 ? i : i - T4 / 8 + T4P / 8];
    for (int i = threadIdx.x; i < 4 * HID + NOUT; i += blockDim.x) {
        int off = i < HID ? B0 + i : i < 2 * HID ? B1 + i - HID : i < 3 * HID ? B2 + i - 2 * HID
                : i < 4 * HID ? B3 + i - 3 * HID : B4 + out_orig(i - 4 * HID);
        sm.bias[i] = params[off];
    }
}

// acc[nt] += Wt (rows nt 16 ..) x X^T over K, X the wave's 16 rows (ldx)
template <int K, int NT>
__device__ __forceinline__ void gemm16_rows_nk_t(const uint16_t* X, int ldx, const uint16_t* Wt, int ldw, f4* acc,
                                                 int lane) {
#pragma unroll
    for (int kt = 0; kt < K / 32; kt++) {
        const int kk = kt * 32 + 8 * (lane >> 4);
        const bf8v x = ld_bf8(X + (lane & 15) * ldx + kk);
#pragma unroll
        for (int nt = 0; nt < NT; nt++)
            acc[nt] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ld_bf8(Wt + (nt * 16 + (lane & 15)) * ldw + kk), x,
                                                              acc[nt], 0, 0, 0);
    }
}

// Two values as packed BF16, rounded to nearest even (gfx950's v_cvt_pk_bf16_f32:
// the rounding of f2bf, one instruction per pair)
typedef __bf16 bf2v __attribute__((ext_vector_type(2)));
__device__ __forceinline__ uint32_t bf16x2(float a, float b) {
    const bf2v v = {(__bf16)a, (__bf16)b};
    return __builtin_bit_cast(uint32_t, v);
}

// ReLU outputs of a hidden layer (bias at b) as BF16 into the wave's rows
__device__ __forceinline__ void store_hidden16(const f4* acc, const float* b, uint16_t* h, int lane) {
    const int m = lane & 15, n4 = 4 * (lane >> 4);
#pragma unroll
    for (int nt = 0; nt < NTH; nt++) {
        float v[4];
#pragma unroll
        for (int j = 0; j < 4; j++) v[j] = fmaxf(acc[nt][j] + b[nt * 16 + n4 + j], 0.0f);
        *reinterpret_cast<uint2*>(h + m * LDH + nt * 16 + n4) = make_uint2(bf16x2(v[0], v[1]), bf16x2(v[2], v[3]));
    }
}

// Layer-0 operand of lane group g for input x (the MFMA's 8 values of bins 8 g ..
// 8 g + 7): the BF16 one-blob values of the bins in blob_window, 0 elsewhere, with
// two exponentials (blob_run from the first window bin the lane holds: the window's
// bounds are multiples of 4, so that bin is 8 g or 8 g + 4, and the run starts
// within the window, where no value underflows)
__device__ __forceinline__ bf8v blob_operand16(float x, int g) {
    const int w0 = blob_window(x), s = 8 * g, st = max(s, w0);
    float v[8];
    blob_run<8>(x, st, 1.0f, v);
    float u[8];
#pragma unroll
    for (int j = 0; j < 8; j++) {
        const int b = s + j;
        const float val = st == s ? v[j] : (j >= 4 ? v[j - 4] : 0.0f);  // bin b is run value b - st
        u[j] = (b >= w0 && b < w0 + 16) ? val : 0.0f;
    }
    typedef unsigned u4v __attribute__((ext_vector_type(4)));
    const u4v w = {bf16x2(u[0], u[1]), bf16x2(u[2], u[3]), bf16x2(u[4], u[5]), bf16x2(u[6], u[7])};
    return __builtin_bit_cast(bf8v, w);
}

// Forward pass of the wave's 16 rows (inputs at in16; rows past the valid ones
// are computed and ignored); o as wave_forward8
__device__ __forceinline__ void wave_forward16(const SmemInf16& sm, WaveState16& ws, const float* in16, int lane,
                                               f4 (&o)[NTO]) {
    const int m = lane & 15, g = lane >> 4, n4 = 4 * g;
    f4 acc[NTH];
#pragma unroll
    for (int nt = 0; nt < NTH; nt++) acc[nt] = f4{0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
    for (int kt = 0; kt < NIN; kt++) {  // layer 0: one input (32 bins) per step of 32
        const bf8v xb = blob_operand16(in16[m * NIN + kt], g);
        const int kk = 32 * kt + 8 * g;
#pragma unroll
        for (int nt = 0; nt < NTH; nt++)
            acc[nt] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ld_bf8(sm.wt + T0 + (nt * 16 + m) * LD0 + kk), xb,
                                                              acc[nt], 0, 0, 0);
    }
    store_hidden16(acc, sm.bias, ws.h, lane);
    wave_sync();
#pragma unroll
    for (int l = 1; l < 4; l++) {
#pragma unroll
        for (int nt = 0; nt < NTH; nt++) acc[nt] = f4{0.0f, 0.0f, 0.0f, 0.0f};
        gemm16_rows_nk_t<HID, NTH>(ws.h, LDH, sm.wt + (l == 1 ? T1 : l == 2 ? T2 : T3), LDH, acc, lane);
        wave_sync();  // every lane has read its operand rows before any overwrites them
        store_hidden16(acc, sm.bias + l * HID, ws.h, lane);
        wave_sync();
    }
#pragma unroll
    for (int nt = 0; nt < NTO; nt++) o[nt] = f4{0.0f, 0.0f, 0.0f, 0.0f};
    gemm16_rows_nk_t<HID, NTO>(ws.h, LDH, sm.wt + T4, LDH, o, lane);
#pragma unroll
    for (int nt = 0; nt < NTO; nt++)
#pragma unroll
        for (int j = 0; j < 4; j++) o[nt][j] = o[nt][j] + sm.bias[4 * HID + nt * 16 + n4 + j];
    wave_sync();
}

// BF16 mixtures of up to 64 samples whose inputs are in ws.in (nin valid)
__device__ __forceinline__ void wave_infer16(const SmemInf16& sm, WaveState16& ws, int nin, int lane, float* dst,
                                             size_t stride) {
    for (int q = 0; q < 4 && 16 * q < nin; q++) {
        f4 o[NTO];
        wave_forward16(sm, ws, ws.in + 16 * q * NIN, lane, o);
        wave_decode(o, lane, min(16, nin - 16 * q), dst + 16 * q, stride);
    }
}

__global__ __launch_bounds__(64 * WAVES16) void k_mlp_mixtures16(const uint16_t* gwt, const float* params,
                                                                 const float* in, int n, float* dst, size_t stride) {
    extern __shared__ __align__(16) unsigned char smem_raw[];
    SmemInf16& sm = *reinterpret_cast<SmemInf16*>(smem_raw);
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    WaveState16& ws = sm.w[wave];
    load_weights16(sm, gwt, params);
    __syncthreads();
    for (int c = blockIdx.x * WAVES16 + wave; c * 64 < n; c += gridDim.x * WAVES16) {
        const int nin = min(64, n - c * 64);
        for (int i = lane; i < 64 * NIN; i += 64) ws.in[i] = i < nin * NIN ? in[(size_t)c * 64 * NIN + i] : 0.0f;
        wave_sync();
        wave_infer16(sm, ws, nin, lane, dst + (size_t)c * 64, stride);
    }
}

// BF16 raw outputs (n x NOUT, network order) by the per-wave inference code
__global__ __launch_bounds__(64 * WAVES16) void k_mlp_forward16(const uint16_t* gwt, const float* params,
                                                                const float* in, int n, float* out) {
    extern __shared__ __align__(16) unsigned char smem_raw[];
    SmemInf16& sm = *reinterpret_cast<SmemInf16*>(smem_raw);
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    WaveState16& ws = sm.w[wave];
    load_weights16(sm, gwt, params);
    __syncthreads();
    for (int t = blockIdx.x * WAVES16 + wave; t * 16 < n; t += gridDim.x * WAVES16) {
        const int nr = min(16, n - t * 16);
        for (int i = lane; i < 16 * NIN; i += 64) ws.in[i] = i < nr * NIN ? in[(size_t)t * 16 * NIN + i] : 0.0f;
        wave_sync();
        f4 o[NTO];
        wave_forward16(sm, ws, ws.in, lane, o);
        const int m = lane & 15, n4 = 4 * (lane >> 4);
        if (m < nr)
#pragma unroll
            for (int nt = 0; nt < NTO; nt++)
#pragma unroll
                for (int j = 0; j < 4; j++) out[((size_t)t * 16 + m) * NOUT + out_orig(nt * 16 + n4 + j)] = o[nt][j];
    }
}

// BF16 mixtures of n input vectors, as k_mlp_mixtures8 (the per-bounce BF16 path's
// encode, forward and decode)
__global__ __launch_bounds__(256) void k_mlp_mixtures(const uint16_t* gwt, const float* params, const float* in, int n,
                                                      float* dst, size_t stride) {
    extern __shared__ __align__(16) unsigned char smem_raw[];
    SmemFwd& sm = *reinterpret_cast<SmemFwd*>(smem_raw);
    __shared__ __align__(16) float o_tile[TILE * NOUT];
    load_weights(sm, gwt, params);
    __syncthreads();
    for (int t = blockIdx.x; t * TILE < n; t += gridDim.x) {
        const int nt = min(TILE, n - t * TILE);
        infer_tile(sm, gwt, o_tile, in + (size_t)t * TILE * NIN, nt, dst + (size_t)t * TILE, stride);
    }
}

// What the host code needs of this width, for code templated on the network.
struct Net {
    static constexpr int kHid = HID, kLobes = NLOBES, kOut = NOUT, kParams = NPARAM, kWt = NWT_ALL, kWq = NWQ,
                         kRow = ROW, kTile = TILE, kChunk = CH, kDwGroups = DW_GROUPS;
    static constexpr int kW[5] = {W0, W1, W2, W3, W4}, kB4 = B4;
    static constexpr size_t kSmemFwd = sizeof(SmemFwd), kSmemFwd8 = sizeof(SmemFwd8), kSmemTrain = sizeof(SmemTrain);
    static constexpr int kWaves8 = WAVES8;
    // per-wave BF16 inference (width 64: the weights fit in LDS)
    static constexpr bool kWave16 = WT_LDS;
    static constexpr int kWaves16 = WAVES16;
    static constexpr size_t kSmemInf16 = sizeof(SmemInf16);
    using Smem16 = SmemInf16;
    static constexpr auto mlp_forward16 = &k_mlp_forward16;
    static constexpr auto mlp_mixtures16 = &k_mlp_mixtures16;
    static __device__ __forceinline__ void load16(SmemInf16& sm, const uint16_t* gwt, const float* params) {
        load_weights16(sm, gwt, params);
    }
    static __device__ __forceinline__ void infer16(const SmemInf16& sm, WaveState16& ws, int nin, int lane,
                                                   float* dst, size_t stride) {
        wave_infer16(sm, ws, nin, lane, dst, stride);
    }
    using Smem8 = SmemFwd8;
    static __device__ __forceinline__ void load8(SmemFwd8& sm, const uint8_t* gwt8, const float* params,
                                                 const float* inv8) {
        load_weights8(sm, gwt8, params, inv8);
    }
    static __device__ __forceinline__ void sync8() { wave_sync(); }
    static __device__ __forceinline__ void infer8(const SmemFwd8& sm, WaveState8& ws, int nin, int lane, float* dst,
                                                  size_t stride) {
        wave_infer8(sm, ws, nin, lane, dst, stride);
    }
    static constexpr auto mlp_forward = &k_mlp_forward;
    static constexpr auto mlp_forward8 = &k_mlp_forward8;
    static constexpr auto mlp_mixtures = &k_mlp_mixtures;
    static constexpr auto mlp_mixtures8 = &k_mlp_mixtures8;
    static constexpr auto train_tile = &k_train_tile;
    static constexpr auto train_dw = &k_train_dw;
    static constexpr auto reduce_partials = &k_reduce_partials;
    static constexpr auto weight_partial = &k_weight_partial;
    static constexpr auto weight_scale = &k_weight_scale;
    static constexpr auto adam = &k_adam;
    static constexpr auto pack_weights = &k_pack_weights;
    static constexpr auto pack_weights2 = &k_pack_weights2;
    static constexpr auto quant8 = &k_quant8;
};
