// build_kernels.hip -- transition-matrix construction on the GPU
// (replaces matrix_cache::precalculate_matrices, src/matrix_cache.cpp:113-163).
//
// The kernels write the pool layouts of engine.h: an FP64 copy (c-major) and,
// when requested, FP32 copies (c-major and row-major) of every matrix scaled
// by 2^kMatExp.  Entries are those of
// get_from_parent_fam_size_to_c (matrix_cache.cpp:78-85): row 0 is delta_0,
// rows s >= 1 come from the birth-death process with a = lt/(1+lt) and
// coeff = 1 - 2a; keys with coeff <= 0 or coeff == 1 (saturated, lt == 1, or
// a == 0) keep only (0,0) = 1 (matrix_cache.cpp:144-154, probability.cpp:157).
#include "engine.h"

namespace cafe {

__device__ __forceinline__ bool degenerate_key(double alpha, double coeff)
{
    return !(coeff > 0.0 && coeff != 1.0);
}

// ---------------------------------------------------------------------------
// Recurrence build.  The one-gene generating function of the process is
// G_1(z) = (a + (1-2a) z) / (1 - a z) and row s of the matrix holds the
// coefficients of G_1(z)^s, so
//     P_s(c) = a P_{s-1}(c) + (1-2a) P_{s-1}(c-1) + a P_s(c-1).
// Every term is non-negative, so rounding errors stay relative (~1e-15 per
// step) and there is no cancellation.
//
// One wavefront per matrix.  Lane l owns rows s = l + 64 b.  Columns advance
// one at a time; along s the recurrence y_s = a y_{s-1} + K_s is a first-order
// linear scan, done across lanes with a Kogge-Stone affine scan and carried
// between 64-row blocks.  Stores for a fixed column are contiguous in s; the
// row-major FP32 copy is written every CW columns from registers.
// ---------------------------------------------------------------------------
// Lane moves of a double within the wave (DPP on both 32-bit halves; lanes
// without a source, or outside the row mask, read 0).
constexpr int kDppRowShr = 0x110;       // row_shr:n, n = 1..15 (rows of 16 lanes)
constexpr int kDppWaveShr1 = 0x138;     // wave_shr:1
constexpr int kDppRowBcast15 = 0x142;   // lane 15 of each row into the next row
constexpr int kDppRowBcast31 = 0x143;   // lane 31 into rows 2 and 3

template <int CTRL, int ROWS>
__device__ __forceinline__ double dpp_move(double v)
{
    const unsigned long long u = __builtin_bit_cast(unsigned long long, v);
    const int lo = __builtin_amdgcn_update_dpp(0, (int)(unsigned)u, CTRL, ROWS, 0xF, false);
    const int hi = __builtin_amdgcn_update_dpp(0, (int)(unsigned)(u >> 32), CTRL, ROWS, 0xF, false);
    return __builtin_bit_cast(double, ((unsigned long long)(unsigned)hi << 32) | (unsigned)lo);
}

__device__ __forceinline__ double read_lane(double v, int lane)
{
    const unsigned long long u = __builtin_bit_cast(unsigned long long, v);
    const int lo = __builtin_amdgcn_readlane((int)(unsigned)u, lane);
    const int hi = __builtin_amdgcn_readlane((int)(unsigned)(u >> 32), lane);
    return __builtin_bit_cast(double, ((unsigned long long)(unsigned)hi << 32) | (unsigned)lo);
}

__device__ __forceinline__ double int_pow(double base, int e)
{
    double r = 1.0;
    while (e) {
        if (e & 1) r *= base;
        base *= base;
        e >>= 1;
    }
    return r;
}

// Columns per row-major write group: a lane keeps CW consecutive columns of
// each of its NB rows in registers (64 floats at most) and then writes
// CW * 4 contiguous bytes per row.
template <int NB> struct RowGroup { static constexpr int v = NB >= 32 ? 2 : NB >= 16 ? 4 : NB >= 2 ? 64 / NB : 32; };

// Any of the three pools may be null: the FP64 pool (c-major), the FP32 copy
// (c-major, entries times 2^kMatExp) and its row-major twin.  copies (may be
// null: every FP32 copy) says per key which FP32 copies its users read
// (kCopy* bits); with kCopyColTop alone only the c-major entries of sizes
// s, c < top are written.  With the row-major copy goes its band (band may be
// null): per 16-row tile the k-tiles between its first and last nonzero
// FP32 entry (a lane tracks its rows' first and last nonzero column as it
// writes them; 16 lanes hold a tile).
template <int NB>
__global__ __launch_bounds__(64) void k_build_recurrence(const double* __restrict__ lam_q,
                                                          const double* __restrict__ t_q,
                                                          int first_slot, int N, int Np,
                                                          long long slot_elems,
                                                          double* __restrict__ mt64,
                                                          float* __restrict__ lf32,
                                                          float* __restrict__ mt32,
                                                          const unsigned char* __restrict__ copies, int top,
                                                          int* __restrict__ band, int band_stride)
{
    constexpr int CW = RowGroup<NB>::v;
    const int key = blockIdx.x;
    const int lane = threadIdx.x;
    const double lt = lam_q[key] * t_q[key];
    const double alpha = lt / (1 + lt);
    const double coeff = 1 - 2 * alpha;
    const long long base = (long long)(first_slot + key) * slot_elems;
    const int use = copies ? copies[key] : (kCopyRow | kCopyCol);
    // c-major entries below `col_rows` rows, in columns below `col_cols`
    const int col_rows = (use & kCopyCol) ? N : (use & kCopyColTop) ? (top < N ? top : N) : 0;
    double* M64 = mt64 ? mt64 + base : nullptr;
    float* L32 = lf32 && col_rows > 0 ? lf32 + base : nullptr;
    float* R32 = mt32 && (use & kCopyRow) ? mt32 + base : nullptr;
    int* BT = R32 && band ? band + (long long)(first_slot + key) * band_stride : nullptr;

    if (degenerate_key(alpha, coeff)) {
        // only entry (0, 0) is nonzero
        if (BT)
            for (int t = lane; t < band_stride; t += 64) BT[t] = t == 0 ? band_entry(0, 1) : band_entry(0, 0);
        for (int c = 0; c < N; ++c)
            for (int b = 0; b < NB; ++b) {
                const int s = lane + 64 * b;
                if (s < N) {
                    const double v = (s == 0 && c == 0) ? 1.0 : 0.0;
                    if (M64) M64[(long long)c * Np + s] = v;
                    if (L32 && s < col_rows && c < col_rows) L32[(long long)c * Np + s] = (float)ldexp(v, kMatExp);
                    if (R32) R32[(long long)s * Np + c] = (float)ldexp(v, kMatExp);
                }
            }
        return;
    }

    // a^d for the scan distances d = 1, 2, 4, 8, and per lane l the factors
    // a^(l mod 16 + 1), a^(l mod 32 + 1) (row broadcasts) and a^(l + 1)
    // (carry from the previous 64-row block)
    double apow[4];
    apow[0] = alpha;
    for (int i = 1; i < 4; ++i) apow[i] = apow[i - 1] * apow[i - 1];
    const double a16 = int_pow(alpha, (lane & 15) + 1);
    const double a32 = int_pow(alpha, (lane & 31) + 1);
    const double a_lane = int_pow(alpha, lane + 1);

    double prev[NB];   // P_s(c-1) for the lane's rows
    int first_nz[NB], last_nz[NB];   // the lane's rows' first and last nonzero FP32 column
#pragma unroll
    for (int b = 0; b < NB; ++b) {
        prev[b] = 0.0;
        first_nz[b] = Np;
        last_nz[b] = -1;
    }

    const int n_groups = (N + CW - 1) / CW;
    for (int g = 0; g < n_groups; ++g) {
        float row[NB][CW];   // this group's columns of the lane's rows (FP32 copy)
#pragma unroll
        for (int cc = 0; cc < CW; ++cc) {
            const int c = g * CW + cc;
            if (c >= N) {   // past the last column (wave-uniform): zero padding
#pragma unroll
                for (int b = 0; b < NB; ++b) row[b][cc] = 0.f;
                continue;
            }
            double K[NB];
            // K_s = (1-2a) P_{s-1}(c-1) + a P_s(c-1); row 0 is delta_0 itself
#pragma unroll
            for (int b = 0; b < NB; ++b) {
                double below = dpp_move<kDppWaveShr1, 0xF>(prev[b]);   // lane l-1, 0 at lane 0
                const double from_prev_block = b > 0 ? read_lane(prev[b > 0 ? b - 1 : 0], 63) : 0.0;
                if (lane == 0) below = from_prev_block;
                K[b] = coeff * below + alpha * prev[b];
            }
            if (lane == 0) K[0] = (c == 0) ? 1.0 : 0.0;   // P_0(c) = delta_{c0}

            double carry = 0.0;
#pragma unroll
            for (int b = 0; b < NB; ++b) {
                // y_s = K_s + a y_{s-1}: within rows of 16 lanes (shifts 1..8),
                // then across rows (broadcast of lane 15 into rows 1 and 3, of
                // lane 31 into rows 2 and 3)
                double y = K[b];
                y = fma(apow[0], dpp_move<kDppRowShr + 1, 0xF>(y), y);
                y = fma(apow[1], dpp_move<kDppRowShr + 2, 0xF>(y), y);
                y = fma(apow[2], dpp_move<kDppRowShr + 4, 0xF>(y), y);
                y = fma(apow[3], dpp_move<kDppRowShr + 8, 0xF>(y), y);
                y = fma(a16, dpp_move<kDppRowBcast15, 0xA>(y), y);
                y = fma(a32, dpp_move<kDppRowBcast31, 0xC>(y), y);
                if (b > 0) y += a_lane * carry;
                carry = read_lane(y, 63);
                prev[b] = y;
                const int s = lane + 64 * b;
                const double v = y < 1.0 ? y : 1.0;   // [0,1] clamp, probability.cpp:147
                const float v32 = (float)ldexp(v, kMatExp);
                row[b][cc] = s < N ? v32 : 0.f;
                if (s < N && v32 != 0.f) {
                    first_nz[b] = first_nz[b] < c ? first_nz[b] : c;
                    last_nz[b] = c;
                }
                if (s < N) {
                    if (M64) M64[(long long)c * Np + s] = v;
                    if (L32 && s < col_rows && c < col_rows) L32[(long long)c * Np + s] = v32;
                }
            }
        }
        if (R32) {
            // row s, columns g CW .. g CW + CW - 1: CW * 4 contiguous bytes
            // (aligned: Np and g CW are multiples of CW)
#pragma unroll
            for (int b = 0; b < NB; ++b) {
                const int s = lane + 64 * b;
                if (s >= N) continue;
                float* dst = R32 + (long long)s * Np + g * CW;
                if constexpr (CW % 4 == 0) {
                    typedef float f4 __attribute__((ext_vector_type(4)));
#pragma unroll
                    for (int q = 0; q < CW / 4; ++q)
                        *(f4*)(dst + 4 * q) = f4{row[b][4 * q], row[b][4 * q + 1], row[b][4 * q + 2], row[b][4 * q + 3]};
                } else {
                    typedef float f2 __attribute__((ext_vector_type(2)));
#pragma unroll
                    for (int q = 0; q < CW / 2; ++q) *(f2*)(dst + 2 * q) = f2{row[b][2 * q], row[b][2 * q + 1]};
                }
            }
        }
    }
    if (BT) {
        // tile 4 b + lane / 16 = rows 64 b + 16 (lane / 16) .. + 15
#pragma unroll
        for (int b = 0; b < NB; ++b) {
            int lo = first_nz[b], hi = last_nz[b];
#pragma unroll
            for (int d = 1; d < 16; d <<= 1) {
                const int olo = __shfl_xor(lo, d);
                const int ohi = __shfl_xor(hi, d);
                lo = lo < olo ? lo : olo;
                hi = hi > ohi ? hi : ohi;
            }
            const int t = 4 * b + (lane >> 4);
            if ((lane & 15) == 0 && t < band_stride)
                BT[t] = hi >= 0 ? band_entry(lo / kTile, hi / kTile + 1) : band_entry(0, 0);
        }
    }
}

// ---------------------------------------------------------------------------
// Closed-form build: probability.cpp:104-148 term by term, one thread per
// (s, c).  chooseln uses the host's glibc log-gamma table so the log-binomials
// are the reference's values; terms are exp(t) * pow(coeff, j), added in
// ascending j with separately rounded products and sums.
// ---------------------------------------------------------------------------
__device__ __forceinline__ double chooseln_dev(const double* __restrict__ lg, int n, int r)
{
    if (r == 0) return 0.0;
    if (n <= 0 || r <= 0) return -__builtin_huge_val();
    return __dsub_rn(__dsub_rn(lg[n + 1], lg[r + 1]), lg[n - r + 1]);
}

__global__ __launch_bounds__(256) void k_build_closed_form(const double* __restrict__ lam_q,
                                                           const double* __restrict__ t_q,
                                                           const double* __restrict__ lg,
                                                           int first_slot, int N, int Np,
                                                           long long slot_elems,
                                                           double* __restrict__ mt64,
                                                           float* __restrict__ mt32,
                                                           float* __restrict__ lf32,
                                                           const unsigned char* __restrict__ copies, int top,
                                                           int* __restrict__ band, int band_stride)
{
    const int key = blockIdx.z;
    const int c = blockIdx.y;
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    const int use = copies ? copies[key] : (kCopyRow | kCopyCol);
    // no band scan here: every k-tile of every row tile counts as nonzero
    if (band && mt32 && (use & kCopyRow) && c == 0 && blockIdx.x == 0)
        for (int t = threadIdx.x; t < band_stride; t += blockDim.x)
            band[(long long)(first_slot + key) * band_stride + t] = band_entry(0, band_stride);
    if (s >= N) return;
    const double lt = lam_q[key] * t_q[key];
    const double alpha = lt / (1 + lt);
    const double coeff = 1 - 2 * alpha;
    double v;
    if (s == 0) {
        v = (c == 0) ? 1.0 : 0.0;
    } else if (degenerate_key(alpha, coeff)) {
        v = 0.0;
    } else {
        const double log_alpha = log(alpha);
        const int m = c < s ? c : s;
        double acc = 0.0;
        for (int j = 0; j <= m; ++j) {
            const double t = __dadd_rn(__dadd_rn(chooseln_dev(lg, s, j),
                                                 chooseln_dev(lg, s + c - 1 - j, s - 1)),
                                       __dmul_rn((double)(s + c - 2 * j), log_alpha));
            const double p = __dmul_rn(exp(t), pow(coeff, (double)j));
            acc = __dadd_rn(acc, p);
        }
        v = acc > 1.0 ? 1.0 : (acc < 0.0 ? 0.0 : acc);
    }
    double* M64 = mt64 + (long long)(first_slot + key) * slot_elems;
    M64[(long long)c * Np + s] = v;
    const bool col = (use & kCopyCol) || ((use & kCopyColTop) && s < top && c < top);
    if (mt32 && (use & kCopyRow))
        mt32[(long long)(first_slot + key) * slot_elems + (long long)s * Np + c] = (float)ldexp(v, kMatExp);
    if (lf32 && col) lf32[(long long)(first_slot + key) * slot_elems + (long long)c * Np + s] = (float)ldexp(v, kMatExp);
}

// ---------------------------------------------------------------------------
// Leaf error model folded into the matrices.  A leaf with observed count k
// contributes sum_d w_d(k) M[s][k-h+d] (probability.cpp:187-198 builds the
// leaf vector, matrix::multiply sums it in ascending column order, and the
// columns outside k-h..k+h add exact zeros), so column k of the folded matrix
//     M'[s][k] = sum_{d, k-h+d >= 0} w_d(k) M[s][k-h+d]      (ascending d)
// is that leaf factor.  FP64: c-major, one block per (key, column), separately
// rounded products and sums in the reference's order; FP32: the same from the
// c-major FP32 copy.  Columns without an error-model row are zero.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_fold_err64(const double* __restrict__ mt64, double* __restrict__ out,
                                                    const double* __restrict__ err, int n_dev, int n_rows, int N,
                                                    int Np, long long slot_elems, int first_slot)
{
    const int c = blockIdx.x;
    const long long base = (long long)(first_slot + blockIdx.y) * slot_elems;
    const double* M = mt64 + base;
    double* O = out + base;
    const int h = (n_dev - 1) / 2;
    for (int s = threadIdx.x; s < Np; s += blockDim.x) {
        double x = 0.0;
        if (c < n_rows && c < N)
            for (int d = 0; d < n_dev; ++d) {
                const int cc = c - h + d;
                if (cc < 0 || cc >= N) continue;
                x = __dadd_rn(x, __dmul_rn(M[(long long)cc * Np + s], err[(long long)c * n_dev + d]));
            }
        O[(long long)c * Np + s] = x;
    }
}

__global__ __launch_bounds__(256) void k_fold_err32(const float* __restrict__ lf32, float* __restrict__ out,
                                                    const float* __restrict__ err, int n_dev, int n_rows, int N,
                                                    int Np, long long slot_elems, int first_slot)
{
    const int c = blockIdx.x;
    const long long base = (long long)(first_slot + blockIdx.y) * slot_elems;
    const float* M = lf32 + base;
    float* O = out + base;
    const int h = (n_dev - 1) / 2;
    for (int s = threadIdx.x; s < Np; s += blockDim.x) {
        float x = 0.f;
        if (c < n_rows && c < N)
            for (int d = 0; d < n_dev; ++d) {
                const int cc = c - h + d;
                if (cc < 0 || cc >= N) continue;
                x += M[(long long)cc * Np + s] * err[(long long)c * n_dev + d];
            }
        O[(long long)c * Np + s] = x;
    }
}

// Zero-fill (used instead of hipMemsetAsync so the calls can be captured).
__global__ void k_fill_zero_u32(uint32_t* __restrict__ p, long long n)
{
    long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    const long long stride = (long long)gridDim.x * blockDim.x;
    for (; i < n; i += stride) p[i] = 0u;
}

}  // namespace cafe

// Host-side launchers (called from api.cpp).
namespace cafe {

// Matrices of keys 0 .. n-1 into slots first_slot ..: the closed form (algo
// 0) writes the FP64 pool and the FP32 copies copies[key] names (all of them
// without copies); the recurrence also skips the pools passed as null.  Where
// the row-major FP32 copy is written, so is its band (Np / 16 entries per
// slot; band may be null).
hipError_t launch_build(int algo, const double* d_lam, const double* d_t, const double* d_lg,
                        int n, int first_slot, int N, int Np, long long slot_elems, double* mt64,
                        float* mt32, float* lf32, hipStream_t stream, const unsigned char* copies, int top,
                        int* band)
{
    if (n <= 0) return hipSuccess;
    const int band_stride = Np / kTile;
    if (algo == 0) {
        // grid.z holds the key index: launch at most 65535 keys at a time
        for (int k0 = 0; k0 < n; k0 += 65535) {
            const int nk = n - k0 < 65535 ? n - k0 : 65535;
            dim3 grid((N + 255) / 256, N, nk);
            hipLaunchKernelGGL(k_build_closed_form, grid, dim3(256), 0, stream, d_lam + k0, d_t + k0,
                               d_lg, first_slot + k0, N, Np, slot_elems, mt64, mt32, lf32,
                               copies ? copies + k0 : nullptr, top, band, band_stride);
        }
    } else {
        if (Np % 32 != 0) return hipErrorInvalidValue;   // row groups of up to 32 columns
        const int nb = (N + 63) / 64;
#define CAFE_BUILD_NB(NB)                                                                                 \
    if (nb <= NB) {                                                                                       \
        hipLaunchKernelGGL(k_build_recurrence<NB>, dim3(n), dim3(64), 0, stream, d_lam, d_t, first_slot, N, \
                           Np, slot_elems, mt64, lf32, mt32, copies, top, band, band_stride);             \
        return hipGetLastError();                                                                         \
    }
        CAFE_BUILD_NB(1)
        CAFE_BUILD_NB(2)
        CAFE_BUILD_NB(4)
        CAFE_BUILD_NB(8)
        CAFE_BUILD_NB(16)
        CAFE_BUILD_NB(32)
#undef CAFE_BUILD_NB
        return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

// Folds the error model into slots first_slot .. first_slot+n-1 of the FP64
// and c-major FP32 pools (every column; padding stays zero); a null output
// pool is skipped.
hipError_t launch_fold(const double* mt64, const float* lf32, double* out64, float* out32, const double* err64,
                       const float* err32, int n_dev, int n_rows, int N, int Np, long long slot_elems,
                       int first_slot, int n, hipStream_t stream)
{
    for (int k0 = 0; k0 < n; k0 += 65535) {
        const int nk = n - k0 < 65535 ? n - k0 : 65535;
        if (out64)
            hipLaunchKernelGGL(k_fold_err64, dim3(Np, nk), dim3(256), 0, stream, mt64, out64, err64, n_dev, n_rows,
                               N, Np, slot_elems, first_slot + k0);
        if (out32)
            hipLaunchKernelGGL(k_fold_err32, dim3(Np, nk), dim3(256), 0, stream, lf32, out32, err32, n_dev, n_rows,
                               N, Np, slot_elems, first_slot + k0);
    }
    return hipGetLastError();
}

hipError_t launch_fill_zero(void* p, long long bytes, hipStream_t stream)
{
    const long long words = bytes / 4;
    if (words <= 0) return hipSuccess;
    long long blocks = (words + 255) / 256;
    if (blocks > 4096) blocks = 4096;
    hipLaunchKernelGGL(k_fill_zero_u32, dim3((unsigned)blocks), dim3(256), 0, stream,
                       (uint32_t*)p, words);
    return hipGetLastError();
}

}  // namespace cafe
