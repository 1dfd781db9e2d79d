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
// One wavefront per matrix.  Lane l owns the R consecutive rows s = R l ..
// R l + R - 1 (R = ceil(N / 64)).  Columns advance one at a time; along s the
// recurrence y_s = a y_{s-1} + K_s is a first-order linear scan: each lane
// runs it over its own rows, one Kogge-Stone affine scan across the lanes
// carries the lanes' last values, and a fixup adds the carry to each row.
// A lane's stores for a fixed column are R contiguous entries (the wave's are
// contiguous in s); the row-major FP32 copy is written every CW columns from
// registers.
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
// each of its R rows in registers (64 floats at most) and then writes CW * 4
// contiguous bytes per row.
template <int R> struct RowGroup { static constexpr int v = R >= 2 ? 64 / R : 32; };

// Any of the three pools may be null: the FP64 pool (c-major), the FP32 copy
// (c-major, entries times 2^kMatExp) and its row-major twin.  copies (may be
// null: every FP32 copy) says per key which FP32 copies its users read
// (kCopy* bits); with kCopyColTop alone only the c-major entries of sizes
// s, c < top are written.  With the row-major copy goes its band (band may be
// null): per 16-row tile the k-tiles between its first and last nonzero
// FP32 entry (a lane tracks its rows' first and last nonzero column as it
// writes them, per tile; the 16 / R lanes of a tile combine theirs).
template <int R>
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
    constexpr int CW = RowGroup<R>::v;
    const int key = blockIdx.x;
    const int lane = threadIdx.x;
    const int s0 = R * lane;   // the lane's rows s0 .. s0 + R - 1
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
            for (int j = 0; j < R; ++j) {
                const int s = s0 + j;
                if (s < N) {
                    const double v = (s == 0 && c == 0) ? 1.0 : 0.0;
                    if (M64) M64[(long long)c * Np + s] = v;
                    if (L32 && s < col_rows && c < col_rows) L32[(long long)c * Np + s] = (float)ldexp(v, kMatExp);
                    if (R32) R32[(long long)s * Np + c] = (float)ldexp(v, kMatExp);
                }
            }
        return;
    }

    // Along s, y_s = K_s + a y_{s-1}: a lane runs its R rows in order from 0
    // (z_j), then the lanes' last values are scanned with factor A = a^R
    // (Kogge-Stone within rows of 16 lanes, then across rows), and the
    // carry Y from the lanes before adds a^(j+1) Y to row j.  Factors: A^d
    // for the shifts d = 1, 2, 4, 8, and per lane A^(l mod 16 + 1),
    // A^(l mod 32 + 1) (row broadcasts); a^(j+1) for the carry.
    double apj[R];
    apj[0] = alpha;
#pragma unroll
    for (int j = 1; j < R; ++j) apj[j] = apj[j - 1] * alpha;
    const double A = apj[R - 1];
    double Apow[4];
    Apow[0] = A;
    for (int i = 1; i < 4; ++i) Apow[i] = Apow[i - 1] * Apow[i - 1];
    const double A16 = int_pow(A, (lane & 15) + 1);
    const double A32 = int_pow(A, (lane & 31) + 1);

    // 16-row tiles per lane (R > 16) and lanes per tile (R < 16)
    constexpr int TPL = R > kTile ? R / kTile : 1;
    constexpr int LPT = R < kTile ? kTile / R : 1;
    double prev[R];   // P_s(c-1) for the lane's rows
    int first_nz[TPL], last_nz[TPL];   // per tile: the lane's rows' first and last nonzero FP32 column
#pragma unroll
    for (int j = 0; j < R; ++j) prev[j] = 0.0;
#pragma unroll
    for (int q = 0; q < TPL; ++q) {
        first_nz[q] = Np;
        last_nz[q] = -1;
    }
    // whole-lane vector stores of a column: every row of the lane is a
    // matrix row (the rows are 8 R-byte aligned: Np is a multiple of 32)
    const bool full = s0 + R <= N;

    const int n_groups = (N + CW - 1) / CW;
    for (int g = 0; g < n_groups; ++g) {
        float row[R][CW];   // this group's columns of the lane's rows (FP32 copy)
#pragma unroll
        for (int cc = 0; cc < CW; ++cc) {
            const int c = g * CW + cc;
            if (c >= N) {   // past the last column (wave-uniform): zero padding
#pragma unroll
                for (int j = 0; j < R; ++j) row[j][cc] = 0.f;
                continue;
            }
            // K_s = (1-2a) P_{s-1}(c-1) + a P_s(c-1); row 0 is delta_0 itself
            const double below0 = dpp_move<kDppWaveShr1, 0xF>(prev[R - 1]);   // lane l-1's last row, 0 at lane 0
            double z[R];
#pragma unroll
            for (int j = 0; j < R; ++j) {
                const double below = j == 0 ? below0 : prev[j - 1];
                double K = coeff * below + alpha * prev[j];
                if (j == 0 && lane == 0) K = (c == 0) ? 1.0 : 0.0;   // P_0(c) = delta_{c0}
                z[j] = j == 0 ? K : fma(alpha, z[j - 1], K);
            }
            // inclusive scan of the lanes' last values: Y_l = z_{R-1}(l) + A Y_{l-1}
            double Y = z[R - 1];
            Y = fma(Apow[0], dpp_move<kDppRowShr + 1, 0xF>(Y), Y);
            Y = fma(Apow[1], dpp_move<kDppRowShr + 2, 0xF>(Y), Y);
            Y = fma(Apow[2], dpp_move<kDppRowShr + 4, 0xF>(Y), Y);
            Y = fma(Apow[3], dpp_move<kDppRowShr + 8, 0xF>(Y), Y);
            Y = fma(A16, dpp_move<kDppRowBcast15, 0xA>(Y), Y);
            Y = fma(A32, dpp_move<kDppRowBcast31, 0xC>(Y), Y);
            const double carry = dpp_move<kDppWaveShr1, 0xF>(Y);   // Y_{l-1}, 0 at lane 0
            double v64[R];
            float v32[R];
#pragma unroll
            for (int j = 0; j < R; ++j) {
                const double y = fma(apj[j], carry, z[j]);
                prev[j] = y;
                const double v = y < 1.0 ? y : 1.0;   // [0,1] clamp, probability.cpp:147
                v64[j] = v;
                v32[j] = (float)ldexp(v, kMatExp);
                const bool in = s0 + j < N;
                row[j][cc] = in ? v32[j] : 0.f;
                if (in && v32[j] != 0.f) {
                    const int q = j / kTile;
                    first_nz[q] = first_nz[q] < c ? first_nz[q] : c;
                    last_nz[q] = c;
                }
            }
            double* m64 = M64 ? M64 + (long long)c * Np + s0 : nullptr;
            float* l32 = L32 && c < col_rows ? L32 + (long long)c * Np + s0 : nullptr;
            if (full) {
                if (m64) {
                    if constexpr (R % 2 == 0) {
                        typedef double d2 __attribute__((ext_vector_type(2)));
#pragma unroll
                        for (int q = 0; q < R / 2; ++q) *(d2*)(m64 + 2 * q) = d2{v64[2 * q], v64[2 * q + 1]};
                    } else {
                        m64[0] = v64[0];
                    }
                }
                if (l32 && s0 + R <= col_rows) {
                    if constexpr (R % 4 == 0) {
                        typedef float f4 __attribute__((ext_vector_type(4)));
#pragma unroll
                        for (int q = 0; q < R / 4; ++q)
                            *(f4*)(l32 + 4 * q) = f4{v32[4 * q], v32[4 * q + 1], v32[4 * q + 2], v32[4 * q + 3]};
                    } else {
#pragma unroll
                        for (int j = 0; j < R; ++j) l32[j] = v32[j];
                    }
                } else if (l32) {
#pragma unroll
                    for (int j = 0; j < R; ++j)
                        if (s0 + j < col_rows) l32[j] = v32[j];
                }
            } else {
#pragma unroll
                for (int j = 0; j < R; ++j) {
                    if (s0 + j >= N) break;
                    if (m64) m64[j] = v64[j];
                    if (l32 && s0 + j < col_rows) l32[j] = v32[j];
                }
            }
        }
        if (R32) {
            // row s, columns g CW .. g CW + CW - 1: CW * 4 contiguous bytes
            // (aligned: Np and g CW are multiples of CW)
#pragma unroll
            for (int j = 0; j < R; ++j) {
                const int s = s0 + j;
                if (s >= N) break;
                float* dst = R32 + (long long)s * Np + g * CW;
                if constexpr (CW % 4 == 0) {
                    typedef float f4 __attribute__((ext_vector_type(4)));
#pragma unroll
                    for (int q = 0; q < CW / 4; ++q)
                        *(f4*)(dst + 4 * q) = f4{row[j][4 * q], row[j][4 * q + 1], row[j][4 * q + 2], row[j][4 * q + 3]};
                } else {
                    typedef float f2 __attribute__((ext_vector_type(2)));
#pragma unroll
                    for (int q = 0; q < CW / 2; ++q) *(f2*)(dst + 2 * q) = f2{row[j][2 * q], row[j][2 * q + 1]};
                }
            }
        }
    }
    if (BT) {
        // tile s0 / 16 + q: the LPT lanes from lane - lane % LPT
#pragma unroll
        for (int q = 0; q < TPL; ++q) {
            int lo = first_nz[q], hi = last_nz[q];
#pragma unroll
            for (int d = 1; d < LPT; d <<= 1) {
                const int olo = __shfl_xor(lo, d);
                const int ohi = __shfl_xor(hi, d);
                lo = lo < olo ? lo : olo;
                hi = hi > ohi ? hi : ohi;
            }
            const int t = s0 / kTile + q;
            if (lane % LPT == 0 && t < band_stride)
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
        const int rows = (N + 63) / 64;   // rows per lane
#define CAFE_BUILD_R(R)                                                                                   \
    if (rows <= R) {                                                                                      \
        hipLaunchKernelGGL(k_build_recurrence<R>, dim3(n), dim3(64), 0, stream, d_lam, d_t, first_slot, N,  \
                           Np, slot_elems, mt64, lf32, mt32, copies, top, band, band_stride);             \
        return hipGetLastError();                                                                         \
    }
        CAFE_BUILD_R(1)
        CAFE_BUILD_R(2)
        CAFE_BUILD_R(4)
        CAFE_BUILD_R(8)
        CAFE_BUILD_R(16)
        CAFE_BUILD_R(32)
#undef CAFE_BUILD_R
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
