// prune_f16.hip -- batched Felsenstein pruning with split-FP16 contractions on
// the matrix cores (CAFE_GPU_PATH_BATCHED16).  It replaces the same reference
// code as prune_mfma.hip: inference_prune / compute_node_probability
// (src/core.cpp:134-145, src/probability.cpp:175-330) and the root reductions
// of base_model.cpp:77-95 and gamma_core.cpp:143-165, for many families at once.
//
// Operands.  An FP32 operand x is carried as two FP16 halves
//     x = hi + lo 2^-11 (+ error),   hi = fp16(x),  lo = fp16((x - hi) 2^11),
// and a contraction X = M v as three v_mfma_f32_16x16x32_f16 products with FP32
// accumulation, X = Mhi vhi + (Mhi vlo + Mlo vhi) 2^-11; the dropped lo lo term
// is below 2^-22 of the result, so in-range values keep about FP32 accuracy.
// Matrices are stored x 2^kMatExp16 (entries <= 1) and partial vectors are
// renormalised per family to a maximum just below 2^vec_exp16, so both stay
// inside FP16's range.
//
// Lane layout.  v_mfma_f32_16x16x32_f16: lane l = 16 g + f holds A row f and B
// column f at k = 8 g + j (j = 0..7); D is the layout of the FP32 kernel (rows
// 4 g + r of column f).  Contraction index x of k-block kb maps to k so that
// lane group g holds x = 32 kb + 4 g + r (j = r) and 32 kb + 16 + 4 g + r
// (j = 4 + r): exactly its own stored rows of the vector's 16-row tiles 2 kb
// and 2 kb + 1 (the fragment layout of prune_mfma.hip).  The FP16 matrix pools
// store each row's 32-column blocks in that order (column 32 kb + 16 h + 4 g + r
// at position 32 kb + 8 g + 4 h + r), so a lane's A operand of one block is one
// 16-byte load.
//
// Flush bound.  FP16's range is narrow, and a scalar bound relative to each
// factor's scale (the FP32 kernel's) would reject nearly every family.  Here a
// partial vector v carries a per-entry bound e >= |v - v_exact| on what the
// narrow range can lose (the rounding of in-range values is not bounded, as in
// the FP32 path), propagated along with v:
//   - representation of v: an entry below 2^-14 (FP16 subnormal for hi) is
//     off by at most 2^-36; the contraction's input bound is
//     ê = e + 2^-36 [0 < v < 2^-14];
//   - representation of M: at most 2^-36 per entry below 2^-14, so at most
//     2^-36 sum_x v over a row (sum_x v <= (1 + 2^-10) sum_x vhi + Lp 2^-25;
//     sum_x vhi is one more all-ones MFMA per k-block);
//   - through M: sum_x M ê <= (1 + 2^-9) Mhi ê + 2^-14 sum_x ê, since
//     M <= Mhi (1 + 2^-10) where Mhi is normal and M < 2^-14 elsewhere; the
//     first term is a fourth MFMA whose B operand is ê x 2^30 rounded up, the
//     sum one more MFMA per k-block with an all-ones A operand;
//   - products of factors: e(ab) <= e_a (b + e_b) + a e_b (+ FP32 flushes).
// At the root a family is certified when max_j e_j prior_j is below 1e-9 of
// its likelihood; otherwise (and when its FP64 value would be subnormal) it
// is flagged for the FP64 recompute, like an uncertified FP32 family.
#include "prune_common.h"

namespace cafe {

typedef float floatx4 __attribute__((ext_vector_type(4)));
typedef _Float16 half8 __attribute__((ext_vector_type(8)));

constexpr float kLoScale = 2048.f;        // 2^11: lo = (x - hi) 2^11
constexpr float kLoInv = 1.f / 2048.f;
constexpr float kNormal16 = 0x1p-14f;     // smallest normal FP16
constexpr float kFloor16 = 0x1p-36f;      // representation floor of a split operand
constexpr float kEScale = 0x1p30f;        // ê's scale in the bound MFMA
constexpr float kEInv = 0x1p-30f;

__device__ __forceinline__ floatx4 mfma32(const half8& a, const half8& b, floatx4 c)
{
    return __builtin_amdgcn_mfma_f32_16x16x32_f16(a, b, c, 0, 0, 0);
}

// Waits for this wave's LDS traffic and meets the workgroup at a barrier,
// leaving global loads in flight (the next step's prefetched operands).
__device__ __forceinline__ void lds_barrier16()
{
    asm volatile("s_waitcnt lgkmcnt(0)\n\ts_barrier" ::: "memory");
}

// Makes the wave wait for a loaded value without using it.
__device__ __forceinline__ void await16(const floatx4& v)
{
#pragma unroll
    for (int j = 0; j < 4; ++j) asm volatile("" ::"v"(v[j]));
}

constexpr int kSwizzleXor16b = 0x401F;   // ds_swizzle: lane l reads lane l ^ 16

// Maximum over the family's four lanes f, f+16, f+32, f+48.
__device__ __forceinline__ float max4(float v)
{
    v = fmaxf(v, __builtin_bit_cast(float, __builtin_amdgcn_ds_swizzle(__builtin_bit_cast(int, v), kSwizzleXor16b)));
    const unsigned u = __builtin_bit_cast(unsigned, v);
    const auto sw = __builtin_amdgcn_permlane32_swap(u, u, false, false);
    return fmaxf(__builtin_bit_cast(float, (unsigned)sw[0]), __builtin_bit_cast(float, (unsigned)sw[1]));
}

// Root reduction with the bound: the best key and the largest e_j prior_j.
struct RootBest16 {
    RootBest b;
    double emax;
};

__device__ __forceinline__ RootBest16 root_combine16(RootBest16 r)
{
    for (int d = 16; d <= 32; d <<= 1) {
        const double ok = __shfl_xor(r.b.key, d);
        const double ox = __shfl_xor(r.b.x, d);
        const int oj = __shfl_xor(r.b.j, d);
        const double om = __shfl_xor(r.b.xmax, d);
        const double oe = __shfl_xor(r.emax, d);
        root_merge(r.b, ok, ox, oj, om);
        r.emax = r.emax > oe ? r.emax : oe;
    }
    return r;
}

// ---------------------------------------------------------------------------
// The kernel: the structure of k_prune_mfma<float, ...> (child steps walked
// with the next step's operands prefetched into the other half of a
// ping-pong buffer; wave wv computes the contiguous run [wv n / W,
// (wv + 1) n / W) of the n output tiles, and an interior step multiplies the
// matrix only in the 32-row k-blocks that meet the union of its rows' bands:
// the FP16 halves of an FP32 zero are zero, so the skipped products add
// exactly 0.  The all-ones sums over the child's blocks, which the bound
// needs, still run over every block, in order).
//
// Partial vectors are stored ready for use as B operands: per 32-row k-block
// and lane, three half8 -- hi, lo and the bound ê x 2^30 rounded up -- of the
// renormalised vector (layout [part][slot][block][lane], in LDS (LDSP) or the
// workgroup's global scratch).  A node is finished when it is stored: the
// waves exchange their per-family maxima (one barrier), each converts its own
// tiles, and the stores are made visible by a second barrier.  The exchange
// also orders every wave's reads of the children before any store, so a
// node's output may take one of its children's slots (the per-lane kernel's
// program, fewer live slots).  The reader only loads and multiplies.
//
// Prefetch of a step: two 16-byte loads per output tile -- an interior step's
// A operand of k-block 0 (hi, lo), or a leaf step's FP32 column (the same
// load twice) -- so both kinds leave the same number of loads in flight.
// ---------------------------------------------------------------------------
typedef _Float16 half4 __attribute__((ext_vector_type(4)));

template <int NTW, int W, bool LDSP>
__global__ __launch_bounds__(64 * W) void k_prune_f16(PruneParams p)
{
    const int lane = threadIdx.x & 63;
    const int wv = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
    const int fl = lane & 15;
    const int g = lane >> 4;
    const _Float16* __restrict__ poolh = p.mt16h;
    const _Float16* __restrict__ pooll = p.mt16l;
    const float* __restrict__ lpool = p.leaf32;
    const int Np = p.Np;
    const int n_slots = p.n_slots;
    const int slot_blocks = (p.slot_tiles + 1) >> 1;
    const int a_exp = p.vec_exp16;
    const int n_steps = p.n_steps;
    const int zero = __builtin_amdgcn_readfirstlane(p.zero);
    const float leaf_e = (float)((p.n_dev > 1 ? p.n_dev : 1) * kTau32);
    constexpr float kTau = 1.1754943508222875e-38f;
    half8 ones;
#pragma unroll
    for (int j = 0; j < 8; ++j) ones[j] = (_Float16)1.f;

    // dynamic LDS: the node's per-wave maxima, per slot the stored vector's
    // exponent (true value = stored 2^E) per family, the tile's leaf counts,
    // then LDS partials
    extern __shared__ __align__(32) unsigned char s_dyn[];
    float* m_wmax = (float*)s_dyn;
    int* m_esum = (int*)(m_wmax + W * 16);
    const int meta = f16_meta_bytes(n_slots, W);
    int* s_cnt = (int*)(s_dyn + meta);
    const long long part = (long long)n_slots * slot_blocks * 64;   // half8 per part
    half8* fh = LDSP ? (half8*)(s_dyn + meta + mfma_count_bytes(p.n_leaves))
                     : (half8*)p.scratch + (long long)blockIdx.x * 3 * part;
    half8* flo = fh + part;
    half8* fe = flo + part;
    __shared__ double s_rkey[W][16], s_rx[W][16], s_rxmax[W][16], s_remax[W][16];
    __shared__ int s_rj[W][16];

    const int n_tiles = *as_const(p.n_tiles_dev);
    const int groups = (gridDim.x % 8 == 0) ? 8 : 1;
    const int per_group = gridDim.x / groups;
    const int grp = blockIdx.x % groups;
    const int t_lo = (int)((long long)n_tiles * grp / groups);
    const int t_hi = (int)((long long)n_tiles * (grp + 1) / groups);
    for (int tile = t_lo + (int)blockIdx.x / groups; tile < t_hi; tile += per_group) {
        const TileDesc td = ld_const(p.tiles, tile);
        const BucketDev bk = ld_const(p.buckets, td.bucket);
        const int nf = td.nf < 0 ? -td.nf : td.nf;
        for (int i = threadIdx.x; i < p.n_leaves * 16; i += 64 * W) {
            const int f = i & 15;
            const int ff = f < nf ? f : 0;
            const int fam = td.nf < 0 ? p.fam_list[td.fam0 + ff] : bk.fam_begin + td.fam0 + ff;
            s_cnt[i] = p.counts_T[(long long)(i >> 4) * p.F + fam];
        }
        __syncthreads();
        const bool valid = fl < nf;
        const int fi = valid ? fl : 0;
        const int fs = td.nf < 0 ? p.fam_list[td.fam0 + fi] : bk.fam_begin + td.fam0 + fi;
        const int dst = p.perm[fs];
        const __attribute__((address_space(4))) int* nslot = as_const(p.node_slot + (long long)td.point * p.n_nodes);
        const int nts = bk.Lp / kTile;           // 16-row tiles of every child vector
        const int nkb = (nts + 1) >> 1;          // 32-row k-blocks
        const int rows_root = bk.Rp / kTile;

        int offA0[NTW], offA1[NTW], offL0[NTW], offL1[NTW];
        // this wave's run of row tiles, non-root [c0, c1) / root [r0, r1)
        const int c0 = wv * nts / W, c1 = (wv + 1) * nts / W;
        const int r0 = wv * rows_root / W, r1 = (wv + 1) * rows_root / W;
#pragma unroll
        for (int i = 0; i < NTW; ++i) {
            const int t0 = c0 + i < c1 ? c0 + i : nts - 1;
            const int t1 = r0 + i < r1 ? r0 + i : rows_root - 1;
            offA0[i] = (kTile * t0 + fl) * Np + 8 * g;
            offA1[i] = (kTile * t1 + fl) * Np + 8 * g;
            offL0[i] = kTile * t0 + 4 * g;
            offL1[i] = kTile * t1 + 4 * g;
        }

        // k-blocks [kbl, kbh) of an interior step: those that meet the union
        // of the bands of the wave's live row tiles (at least one block; a
        // one-wave workgroup holds every row tile: no lookup)
        const __attribute__((address_space(4))) int* band = as_const(p.band);
        auto kband = [&](int slot, bool root, int& kbl, int& kbh) {
            kbl = 0;
            kbh = nkb;
            if constexpr (W > 1) {
                if (!band) return;
                const int a = root ? r0 : c0, b = root ? r1 : c1;
                int lo = nts, hi = 0;
#pragma unroll
                for (int i = 0; i < NTW; ++i)
                    if (a + i < b) {
                        const int e = band[(long long)slot * p.band_stride + a + i];
                        lo = lo < (e & 0xffff) ? lo : (e & 0xffff);
                        hi = hi > (e >> 16) ? hi : (e >> 16);
                    }
                int bl = lo >> 1, bh = (hi + 1) >> 1;
                bh = bh < nkb ? bh : nkb;
                if (bl >= bh) {
                    bl = 0;
                    bh = 1;
                }
                kbl = bl;
                kbh = bh;
            }
        };
        auto prefetch = [&](floatx4 (&pf)[2][NTW], StepDev& sd, int& kbl, int& kbh, int st) {
            st = __builtin_amdgcn_readfirstlane(st < n_steps ? st : n_steps - 1);
            sd = ld_const(p.steps, st);
            const bool root = (sd.ctl & kStepRoot) != 0;
            const bool leaf = sd.leaf_col >= 0;
            const int slot = nslot[sd.node];
            const long long mo = (long long)slot * p.slot_elems;
            kbl = 0;
            kbh = nkb;
            if (!leaf) kband(slot, root, kbl, kbh);
            const int cn = s_cnt[(leaf ? sd.leaf_col : 0) * 16 + fl] * Np;
#pragma unroll
            for (int i = 0; i < NTW; ++i) {
                const int oa = root ? offA1[i] : offA0[i];
                const int ol = cn + (root ? offL1[i] : offL0[i]);
                if (leaf) {
                    const floatx4* q = (const floatx4*)(lpool + mo + ol);
                    pf[0][i] = q[0];
                    pf[1][i] = q[zero];
                } else {
                    pf[0][i] = *(const floatx4*)(poolh + mo + oa + 32 * kbl);
                    pf[1][i] = *(const floatx4*)(pooll + mo + oa + 32 * kbl);
                }
            }
        };

        floatx4 acc[NTW], acc_e[NTW];
        int e_sum = 0;

        auto consume = [&](const floatx4 (&pf)[2][NTW], const StepDev& sd, int kbl, int kbh) {
            const bool root = (sd.ctl & kStepRoot) != 0;
            const int rows = root ? rows_root : nts;
            const int out_slot = sd.ctl & kStepSlotMask;
            int tt[NTW];
            bool live[NTW];
#pragma unroll
            for (int i = 0; i < NTW; ++i) {
                const int u = (root ? r0 : c0) + i;
                live[i] = u < (root ? r1 : c1);
                tt[i] = live[i] ? u : rows - 1;
            }
            const bool first = (sd.ctl & kStepFirst) != 0;
            if (first) e_sum = 0;
            floatx4 x[NTW], xe[NTW];
            if (sd.leaf_col >= 0) {
                e_sum -= kMatExp;
#pragma unroll
                for (int i = 0; i < NTW; ++i) {
                    x[i] = pf[0][i];
                    await16(pf[1][i]);
                    xe[i] = floatx4{leaf_e, leaf_e, leaf_e, leaf_e};
                }
            } else {
                e_sum += m_esum[sd.in_slot * 16 + fl] - kMatExp16;
                const long long mo = (long long)nslot[sd.node] * p.slot_elems;
                const _Float16* __restrict__ Mh = poolh + mo;
                const _Float16* __restrict__ Ml = pooll + mo;
                const long long b0 = (long long)sd.in_slot * slot_blocks * 64 + lane;
                floatx4 xh[NTW], xl[NTW], xb[NTW];
                floatx4 xs = floatx4{0, 0, 0, 0};   // sum_x ê x 2^30, every row
                floatx4 xv = floatx4{0, 0, 0, 0};   // sum_x vhi, every row
#pragma unroll
                for (int i = 0; i < NTW; ++i) {
                    xh[i] = floatx4{0, 0, 0, 0};
                    xl[i] = floatx4{0, 0, 0, 0};
                    xb[i] = floatx4{0, 0, 0, 0};
                }
                int ao[NTW];
#pragma unroll
                for (int i = 0; i < NTW; ++i) ao[i] = root ? offA1[i] : offA0[i];
                auto block = [&](const floatx4 (&ah)[NTW], const floatx4 (&al)[NTW], const half8& bh,
                                 const half8& bl, const half8& be) {
#pragma unroll
                    for (int i = 0; i < NTW; ++i) {
                        const half8 a_h = __builtin_bit_cast(half8, ah[i]);
                        const half8 a_l = __builtin_bit_cast(half8, al[i]);
                        xh[i] = mfma32(a_h, bh, xh[i]);
                        xl[i] = mfma32(a_h, bl, xl[i]);
                        xl[i] = mfma32(a_l, bh, xl[i]);
                        xb[i] = mfma32(a_h, be, xb[i]);
                    }
                    xs = mfma32(ones, be, xs);   // sums over the block of ê and of vhi
                    xv = mfma32(ones, bh, xv);
                };
                // blocks below the band: the sums only
                for (int kb = 0; kb < kbl; ++kb) {
                    xs = mfma32(ones, fe[b0 + kb * 64], xs);
                    xv = mfma32(ones, fh[b0 + kb * 64], xv);
                }
                block(pf[0], pf[1], fh[b0 + kbl * 64], flo[b0 + kbl * 64], fe[b0 + kbl * 64]);
                if (kbh - kbl > 1) {
                    // blocks kbl + 1 .. kbh - 1 through a ring of RD A and B
                    // operand sets (one for eight-wave shapes: registers)
                    constexpr int RD = W >= 8 ? 1 : 2;
                    floatx4 rh[RD][NTW], rl[RD][NTW];
                    half8 qh[RD], ql[RD], qe[RD];
#pragma unroll
                    for (int d = 0; d < RD; ++d) {
                        const int kb = kbl + 1 + d < kbh ? kbl + 1 + d : kbh - 1;
#pragma unroll
                        for (int i = 0; i < NTW; ++i) {
                            rh[d][i] = *(const floatx4*)(Mh + ao[i] + 32 * kb);
                            rl[d][i] = *(const floatx4*)(Ml + ao[i] + 32 * kb);
                        }
                        qh[d] = fh[b0 + kb * 64];
                        ql[d] = flo[b0 + kb * 64];
                        qe[d] = fe[b0 + kb * 64];
                    }
                    for (int kb = kbl + 1; kb < kbh; kb += RD) {
#pragma unroll
                        for (int d = 0; d < RD; ++d) {
                            const int k = kb + d;
                            if (k < kbh) block(rh[d], rl[d], qh[d], ql[d], qe[d]);
                            const int kn = k + RD < kbh ? k + RD : kbh - 1;
#pragma unroll
                            for (int i = 0; i < NTW; ++i) {
                                rh[d][i] = *(const floatx4*)(Mh + ao[i] + 32 * kn);
                                rl[d][i] = *(const floatx4*)(Ml + ao[i] + 32 * kn);
                            }
                            qh[d] = fh[b0 + kn * 64];
                            ql[d] = flo[b0 + kn * 64];
                            qe[d] = fe[b0 + kn * 64];
                        }
                    }
                    // the ring's last loads are never consumed: take them
#pragma unroll
                    for (int d = 0; d < RD; ++d) {
#pragma unroll
                        for (int i = 0; i < NTW; ++i) {
                            await16(rh[d][i]);
                            await16(rl[d][i]);
                        }
                        await16(__builtin_bit_cast(floatx4, qh[d]));
                        await16(__builtin_bit_cast(floatx4, ql[d]));
                        await16(__builtin_bit_cast(floatx4, qe[d]));
                    }
                }
                // blocks above the band: the sums only
                for (int kb = kbh; kb < nkb; ++kb) {
                    xs = mfma32(ones, fe[b0 + kb * 64], xs);
                    xv = mfma32(ones, fh[b0 + kb * 64], xv);
                }
                const float s_e = xs[0] * 0x1p-14f;
                // the matrix's representation floor over a row
                const float floorA = kFloor16 * (xv[0] * (1.f + 0x1p-9f) + (float)bk.Lp * 0x1p-25f);
#pragma unroll
                for (int i = 0; i < NTW; ++i) {
                    x[i] = xh[i] + xl[i] * kLoInv;
#pragma unroll
                    for (int r = 0; r < 4; ++r) xe[i][r] = (xb[i][r] + s_e) * (kEInv * (1.f + 0x1p-9f)) + floorA;
                }
            }
            if (first) {
#pragma unroll
                for (int i = 0; i < NTW; ++i) {
                    acc[i] = x[i];
                    acc_e[i] = xe[i];
                }
            } else {
#pragma unroll
                for (int i = 0; i < NTW; ++i)
#pragma unroll
                    for (int r = 0; r < 4; ++r) {
                        acc_e[i][r] = fmaf(acc_e[i][r], x[i][r] + xe[i][r], fmaf(acc[i][r], xe[i][r], 2.f * kTau));
                        acc[i][r] *= x[i][r];
                    }
            }
            if (!(sd.ctl & kStepLast)) return;

            if (!root) {
                // finish the node: per-family maximum over every wave's rows
                // (rows past the bound are 0), shift to just below 2^a, split
                float wmax = 0.f;
#pragma unroll
                for (int i = 0; i < NTW; ++i) {
                    if (!live[i]) continue;
#pragma unroll
                    for (int r = 0; r < 4; ++r) {
                        if (kTile * tt[i] + 4 * g + r >= bk.L) {
                            acc[i][r] = 0.f;
                            acc_e[i][r] = 0.f;
                        }
                        wmax = fmaxf(wmax, acc[i][r]);
                    }
                }
                float m = max4(wmax);
                if (W > 1) {
                    if (g == 0) m_wmax[wv * 16 + fl] = m;
                    if (!LDSP) __syncthreads();
                    else lds_barrier16();
                    m = 0.f;
#pragma unroll
                    for (int w2 = 0; w2 < W; ++w2) m = fmaxf(m, m_wmax[w2 * 16 + fl]);
                }
                int ex = 0;
                if (m > 0.f) (void)frexpf(m, &ex);
                const int bsh = a_exp - ex;
#pragma unroll
                for (int i = 0; i < NTW; ++i) {
                    if (!live[i]) continue;
                    half4 h4, l4, e4;
#pragma unroll
                    for (int r = 0; r < 4; ++r) {
                        const float xv = ldexpf(acc[i][r], bsh);
                        const _Float16 h = (_Float16)xv;
                        h4[r] = h;
                        l4[r] = (_Float16)((xv - (float)h) * kLoScale);
                        const float ev = ldexpf(acc_e[i][r], bsh) + ((xv > 0.f && xv < kNormal16) ? kFloor16 : 0.f);
                        // (1 + 2^-8) covers the FP16 rounding and the FP32
                        // rounding of the product rule; 2^-24 a subnormal result
                        e4[r] = (_Float16)(ev * (kEScale * (1.f + 0x1p-8f)) + 0x1p-24f);
                    }
                    const long long q = ((long long)out_slot * slot_blocks + (tt[i] >> 1)) * 64 + lane;
                    const int half_ = tt[i] & 1;
                    ((half4*)(fh + q))[half_] = h4;
                    ((half4*)(flo + q))[half_] = l4;
                    ((half4*)(fe + q))[half_] = e4;
                    if (tt[i] == nts - 1 && !half_) {   // odd tile count: the block's second half is 0
                        const half4 z = half4{0, 0, 0, 0};
                        ((half4*)(fh + q))[1] = z;
                        ((half4*)(flo + q))[1] = z;
                        ((half4*)(fe + q))[1] = z;
                    }
                }
                if (wv == 0 && g == 0) m_esum[out_slot * 16 + fl] = e_sum + ex - a_exp;
                if (!LDSP) __syncthreads();
                else if (W > 1) lds_barrier16();
                else asm volatile("" ::: "memory");
                return;
            }

            // root: best x_j prior_j over this lane's rows (row = root size
            // j + 1), the largest e_j prior_j, then across lanes and waves
            const double* __restrict__ pr = p.prior64 + bk.prior_off;
            double pv[NTW][4];
#pragma unroll
            for (int i = 0; i < NTW; ++i)
#pragma unroll
                for (int r = 0; r < 4; ++r) {
                    const int j = kTile * tt[i] + 4 * g + r - 1;
                    pv[i][r] = pr[j < 0 ? 0 : (j < bk.R ? j : bk.R - 1)];
                }
#pragma unroll
            for (int i = 0; i < NTW; ++i)
#pragma unroll
                for (int r = 0; r < 4; ++r) asm volatile("" ::"v"(pv[i][r]));
            RootBest16 rb{{-__builtin_huge_val(), 0.0, -1, 0.0}, 0.0};
#pragma unroll
            for (int i = 0; i < NTW; ++i) {
                if (!live[i]) continue;
#pragma unroll
                for (int r = 0; r < 4; ++r) {
                    const int j = kTile * tt[i] + 4 * g + r - 1;
                    if (j < 0 || j >= bk.R) continue;
                    const double xv = (double)acc[i][r];
                    if (p.root_out && valid) p.root_out[(long long)dst * p.root_stride + j] = ldexp(xv, e_sum);
                    root_take(rb.b, xv * pv[i][r], xv, j);
                    double ek = (double)acc_e[i][r] * pv[i][r];
                    if (!(ek <= 1.0e300)) ek = __builtin_huge_val();   // NaN or overflow: not certified
                    rb.emax = rb.emax > ek ? rb.emax : ek;
                }
            }
            rb = root_combine16(rb);
            if (W > 1) {
                if (g == 0) {
                    s_rkey[wv][fl] = rb.b.key;
                    s_rx[wv][fl] = rb.b.x;
                    s_rxmax[wv][fl] = rb.b.xmax;
                    s_rj[wv][fl] = rb.b.j;
                    s_remax[wv][fl] = rb.emax;
                }
                lds_barrier16();
                if (wv == 0)
                    for (int w2 = 1; w2 < W; ++w2) {
                        root_merge(rb.b, s_rkey[w2][fl], s_rx[w2][fl], s_rj[w2][fl], s_rxmax[w2][fl]);
                        rb.emax = rb.emax > s_remax[w2][fl] ? rb.emax : s_remax[w2][fl];
                    }
            }
            if (wv == 0) {
                double val = 0.0;
                int flag = 0;
                if (!(rb.b.key > 0)) {
                    flag = kFlagRecompute64;
                } else {
                    const double log_scale = (double)e_sum * kLn2;
                    if (p.model == 0)
                        val = log(rb.b.x) + log(p.prior64[bk.prior_off + rb.b.j]) + log_scale;
                    else
                        val = ldexp(rb.b.key, e_sum);
                    // the bound (same units as the key) against the likelihood,
                    // and the FP64 range of the unscaled values
                    if (!(rb.emax * 1.01 <= kCertEta * rb.b.key)) flag |= kFlagRecompute64;
                    if (!(log(rb.b.xmax) + log_scale >= log(kFp64Floor))) flag |= kFlagRecompute64;
                    if (!(log(rb.b.key) + log_scale >= log(kFp64Floor))) flag |= kFlagRecompute64;
                }
                if (valid && g == 0) {
                    p.out_val[(long long)td.point * p.F + dst] = val;
                    p.out_flag[(long long)td.point * p.F + dst] = flag;
                }
            }
        };

        floatx4 pfa[2][NTW], pfb[2][NTW];
        StepDev sda, sdb;
        int kla, kha, klb, khb;
        prefetch(pfa, sda, kla, kha, 0);
        for (int st = 0; st < n_steps; st += 2) {
            prefetch(pfb, sdb, klb, khb, st + 1);
            consume(pfa, sda, kla, kha);
            if (st + 1 < n_steps) {
                prefetch(pfa, sda, kla, kha, st + 2);
                consume(pfb, sdb, klb, khb);
            }
        }
#pragma unroll
        for (int k = 0; k < 2; ++k)
#pragma unroll
            for (int i = 0; i < NTW; ++i) {
                await16(pfa[k][i]);
                await16(pfb[k][i]);
            }
        __syncthreads();
    }
}

// ---------------------------------------------------------------------------
// FP16 pools from the row-major FP32 pool (entries x 2^kMatExp):
// hi = fp16(M 2^kMatExp16), lo = fp16((M 2^kMatExp16 - hi) 2^11), row-major
// with each 32-column block in the lane order above.  One thread per output
// position; rows and columns run over the whole Np x Np slot (padding is 0).
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_split_f16(const float* __restrict__ mt32, _Float16* __restrict__ hi,
                                                    _Float16* __restrict__ lo, int Np, long long slot_elems,
                                                    long long total)
{
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= total) return;
    const long long slot = i / slot_elems;
    const int rem = (int)(i - slot * slot_elems);
    const int row = rem / Np;
    const int pos = rem - row * Np;          // position within the stored row
    const int kb = pos >> 5, q = pos & 31;   // q = 8 g + 4 h + r
    const int gq = q >> 3, h = (q >> 2) & 1, r = q & 3;
    const int col = 32 * kb + 16 * h + 4 * gq + r;
    const float x = ldexpf(mt32[slot * slot_elems + (long long)row * Np + col], kMatExp16 - kMatExp);
    const _Float16 xh = (_Float16)x;
    hi[i] = xh;
    lo[i] = (_Float16)((x - (float)xh) * kLoScale);
}

hipError_t launch_split_f16(const float* mt32, _Float16* hi, _Float16* lo, int Np, long long slot_elems, int n_slots,
                            hipStream_t st)
{
    const long long total = slot_elems * n_slots;
    if (total <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_split_f16, dim3((unsigned)((total + 255) / 256)), dim3(256), 0, st, mt32, hi, lo, Np,
                       slot_elems, total);
    return hipGetLastError();
}

template <int NTW, int W, bool LDSP>
static hipError_t f16_one(int grid, size_t shm, const PruneParams& p, hipStream_t st, int* per_cu)
{
    if (per_cu) return hipOccupancyMaxActiveBlocksPerMultiprocessor(per_cu, k_prune_f16<NTW, W, LDSP>, 64 * W, shm);
    hipLaunchKernelGGL((k_prune_f16<NTW, W, LDSP>), dim3(grid), dim3(64 * W), shm, st, p);
    return hipGetLastError();
}

// One instance per shape (the FP32 kernel's shapes up to eight waves: vectors
// of up to 32 tiles); with per_cu set, reports how many workgroups of it fit
// on one CU instead of launching.
hipError_t f16_dispatch(const MfmaShape& s, int grid, const PruneParams& p, hipStream_t st, int* per_cu)
{
#define CAFE_F16_CASE(NTW, W, LDSP) \
    if (s.ntw == NTW && s.w == W && s.lds == LDSP) return f16_one<NTW, W, LDSP>(grid < 1 ? 1 : grid, s.shm, p, st, per_cu);
    CAFE_F16_CASE(1, 1, true) CAFE_F16_CASE(1, 1, false) CAFE_F16_CASE(2, 1, true) CAFE_F16_CASE(2, 1, false)
    CAFE_F16_CASE(3, 1, true) CAFE_F16_CASE(3, 1, false) CAFE_F16_CASE(4, 1, true) CAFE_F16_CASE(4, 1, false)
    CAFE_F16_CASE(3, 2, true) CAFE_F16_CASE(3, 2, false) CAFE_F16_CASE(4, 2, true) CAFE_F16_CASE(4, 2, false)
    CAFE_F16_CASE(3, 4, true) CAFE_F16_CASE(3, 4, false) CAFE_F16_CASE(4, 4, true) CAFE_F16_CASE(4, 4, false)
    CAFE_F16_CASE(3, 8, false) CAFE_F16_CASE(4, 8, false)
#undef CAFE_F16_CASE
    return hipErrorInvalidValue;
}

}  // namespace cafe
