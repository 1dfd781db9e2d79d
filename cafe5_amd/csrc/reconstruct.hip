// reconstruct.hip -- Pupko's joint ancestral reconstruction on the GPU
// (replaces pupko_reconstructor::reconstruct_gene_family and its node steps,
// src/gene_family_reconstructor.cpp:30-189).
//
// One wave per family.  Nodes run children first; for every node the wave
// keeps, per parent size i (lanes over i), the best child-size value
//     L[i] = max_j  (prod_children L_child[j]) * M[i][j]
// and its argmax C[i] (the first j reaching the maximum, as the reference's
// strict ">" scan from j = 0 keeps it).  The product over children is formed
// once per j in the reference's order (1.0 times each child's entry in
// descendant order) and then multiplied by the matrix entry, so every value
// and every comparison is the reference's, bit for bit.  Leaves hold column
// `count` of their matrix.  The root scans j = 1..Rr-1 against the prior
// (root_equilibrium_distribution::compute(j)) and keeps its argmax as the
// root's state; the wave's first lane then walks the tree parents-first and
// reads each interior node's state from C at its parent's state.
//
// The matrix-vector work is max-times, not a sum of products, so it runs on
// the vector units (FP64), with the matrix column j read coalesced across the
// lanes (c-major FP64 pool: entry (i, j) at j * Np + i).
#include <algorithm>

#include "engine.h"

namespace cafe {

// Families fam_sorted[0 .. nf-1] (count-row order), one per workgroup; L, C
// and the children's product in per-family scratch rows of Lp entries.
// err[0]: a count outside the matrix, err[1]: a parent state outside a node's
// rows (plain stores of 1; any family may set them).
__global__ __launch_bounds__(64) void k_reconstruct(const ReconNode* __restrict__ nodes, int n_nodes,
                                                    const int* __restrict__ kids, const int* __restrict__ parent,
                                                    const double* __restrict__ mt64, int N, int Np,
                                                    long long slot_elems, const int* __restrict__ counts_T, int F,
                                                    const int* __restrict__ fam_sorted, int nf,
                                                    const double* __restrict__ prior, int B, int Rr,
                                                    double* __restrict__ L_all, int* __restrict__ C_all,
                                                    double* __restrict__ prod_all, int Lp,
                                                    int* __restrict__ states, int* __restrict__ err)
{
    const int lane = threadIdx.x;
    const int k = blockIdx.x;
    if (k >= nf) return;
    const int fs = fam_sorted[k];
    double* L = L_all + (long long)k * n_nodes * Lp;   // [node][i]
    int* C = C_all + (long long)k * n_nodes * Lp;
    double* prod = prod_all + (long long)k * Lp;
    int root = -1, root_state = 0;

    for (int s = 0; s < n_nodes; ++s) {
        const ReconNode nd = nodes[s];
        double* Lu = L + (long long)nd.node * Lp;
        if (nd.leaf_col >= 0) {
            // gene_family_reconstructor.cpp:30-46: column `count`, i = parent size
            int obs = counts_T[(long long)nd.leaf_col * F + fs];
            if (obs < 0 || obs >= N) {   // outside the matrix: flagged, read as 0
                if (lane == 0) err[0] = 1;
                obs = 0;
            }
            const double* M = mt64 + (long long)nd.slot * slot_elems + (long long)obs * Np;
            for (int i = lane; i <= B; i += 64) {
                Lu[i] = M[i];
                C[(long long)nd.node * Lp + i] = obs;
            }
            __syncthreads();
            continue;
        }
        // the children's product for every child size j (reference order)
        const int jmax = nd.slot < 0 ? Rr - 1 : B;
        for (int j = lane; j <= jmax; j += 64) {
            double v = 1.0;
            for (int c = 0; c < nd.n_kids; ++c) v *= L[(long long)kids[nd.kid0 + c] * Lp + j];
            prod[j] = v;
        }
        __syncthreads();
        if (nd.slot < 0) {
            // :48-74: the best j = 1..Rr-1 of prod[j] * prior(j); every L[i],
            // i >= 1, is that maximum and C[0] its j
            root = nd.node;
            if (lane == 0) {
                double best = -1;
                int arg = 0;
                for (int j = 1; j < Rr; ++j) {
                    const double val = prod[j] * prior[j];
                    if (val > best) {
                        best = val;
                        arg = j;
                    }
                }
                root_state = arg;
                for (int i = 1; i < Rr; ++i) Lu[i] = best;
            }
            __syncthreads();
            continue;
        }
        // :76-111: L[i] = max_j prod[j] * M[i][j], C[i] = its first argmax
        const double* M = mt64 + (long long)nd.slot * slot_elems;
        for (int i = lane; i <= B; i += 64) {
            double best = -1;
            int arg = 0;
            for (int j = 0; j <= B; ++j) {
                const double val = prod[j] * M[(long long)j * Np + i];
                if (val > best) {
                    best = val;
                    arg = j;
                }
            }
            Lu[i] = best;
            C[(long long)nd.node * Lp + i] = arg;
        }
        __syncthreads();
    }
    // :170-189: states parents first (the reverse of the children-first order)
    if (lane == 0) {
        int* st = states + (long long)k * n_nodes;
        for (int s = n_nodes - 1; s >= 0; --s) {
            const ReconNode nd = nodes[s];
            if (nd.node == root) st[nd.node] = root_state;
            else if (nd.leaf_col >= 0) st[nd.node] = C[(long long)nd.node * Lp];
            else {
                int ps = st[parent[nd.node]];
                if (ps < 0 || ps > B) {   // a parent state past this node's rows
                    err[1] = 1;
                    ps = 0;
                }
                st[nd.node] = C[(long long)nd.node * Lp + ps];
            }
        }
    }
}

// The same reconstruction with kRecW families per workgroup (one wave each)
// for vectors of at most 64 * kRecQ entries (B < 256).  Every family of a
// launch uses the same matrices, so the workgroup stages each interior
// branch's matrix in LDS, a chunk of jc child sizes j at a time, and all its
// waves take their max-times over the staged rows: the matrix is read from
// memory once per workgroup instead of once per family.  Within a family the
// comparisons run over j in ascending order with the first strict maximum
// kept, as in k_reconstruct, so the states are the same.  Waves past the
// launch's last family take part in the loads and barriers only.
constexpr int kRecW = 8;
constexpr int kRecQ = 4;

__global__ __launch_bounds__(64 * kRecW) void k_reconstruct_shared(
    const ReconNode* __restrict__ nodes, int n_nodes, const int* __restrict__ kids, const int* __restrict__ parent,
    const double* __restrict__ mt64, int N, int Np, long long slot_elems, const int* __restrict__ counts_T, int F,
    const int* __restrict__ fam_sorted, int nf, const double* __restrict__ prior, int B, int Rr,
    double* __restrict__ L_all, int* __restrict__ C_all, double* __restrict__ prod_all, int Lp,
    int* __restrict__ states, int* __restrict__ err, int jc)
{
    extern __shared__ double sM[];   // [jc][B + 1]: child sizes j0 .. j0+jc-1, parent sizes 0..B
    const int lane = threadIdx.x & 63;
    const int k = blockIdx.x * kRecW + (int)(threadIdx.x >> 6);
    const bool valid = k < nf;
    const int kk = valid ? k : 0;
    const int fs = fam_sorted[kk];
    const int R1 = B + 1;
    double* L = L_all + (long long)kk * n_nodes * Lp;   // [node][i]
    int* C = C_all + (long long)kk * n_nodes * Lp;
    double* prod = prod_all + (long long)kk * Lp;
    int root = -1, root_state = 0;

    for (int s = 0; s < n_nodes; ++s) {
        const ReconNode nd = nodes[s];
        double* Lu = L + (long long)nd.node * Lp;
        if (nd.leaf_col >= 0) {
            if (valid) {
                int obs = counts_T[(long long)nd.leaf_col * F + fs];
                if (obs < 0 || obs >= N) {
                    if (lane == 0) err[0] = 1;
                    obs = 0;
                }
                const double* M = mt64 + (long long)nd.slot * slot_elems + (long long)obs * Np;
                for (int i = lane; i <= B; i += 64) {
                    Lu[i] = M[i];
                    C[(long long)nd.node * Lp + i] = obs;
                }
            }
            __syncthreads();
            continue;
        }
        const int jmax = nd.slot < 0 ? Rr - 1 : B;
        if (valid) {
            for (int j = lane; j <= jmax; j += 64) {
                double v = 1.0;
                for (int c = 0; c < nd.n_kids; ++c) v *= L[(long long)kids[nd.kid0 + c] * Lp + j];
                prod[j] = v;
            }
        }
        __syncthreads();
        if (nd.slot < 0) {
            root = nd.node;
            if (valid && lane == 0) {
                double best = -1;
                int arg = 0;
                for (int j = 1; j < Rr; ++j) {
                    const double val = prod[j] * prior[j];
                    if (val > best) {
                        best = val;
                        arg = j;
                    }
                }
                root_state = arg;
                for (int i = 1; i < Rr; ++i) Lu[i] = best;
            }
            __syncthreads();
            continue;
        }
        const double* M = mt64 + (long long)nd.slot * slot_elems;
        double best[kRecQ];
        int arg[kRecQ];
#pragma unroll
        for (int q = 0; q < kRecQ; ++q) {
            best[q] = -1;
            arg[q] = 0;
        }
        for (int j0 = 0; j0 <= B; j0 += jc) {
            const int jn = min(jc, R1 - j0);
            // rows j0 .. j0+jn-1 of the matrix's parent-size columns 0..B
            for (int t = threadIdx.x; t < jn * R1; t += 64 * kRecW) {
                const int r = t / R1;
                const int i = t - r * R1;
                sM[t] = M[(long long)(j0 + r) * Np + i];
            }
            __syncthreads();
            if (valid) {
                for (int r = 0; r < jn; ++r) {
                    const double pj = prod[j0 + r];
#pragma unroll
                    for (int q = 0; q < kRecQ; ++q) {
                        const int i = lane + 64 * q;
                        if (i <= B) {
                            const double val = pj * sM[r * R1 + i];
                            if (val > best[q]) {
                                best[q] = val;
                                arg[q] = j0 + r;
                            }
                        }
                    }
                }
            }
            __syncthreads();
        }
        if (valid) {
#pragma unroll
            for (int q = 0; q < kRecQ; ++q) {
                const int i = lane + 64 * q;
                if (i <= B) {
                    Lu[i] = best[q];
                    C[(long long)nd.node * Lp + i] = arg[q];
                }
            }
        }
        __syncthreads();
    }
    if (valid && lane == 0) {
        int* st = states + (long long)k * n_nodes;
        for (int s = n_nodes - 1; s >= 0; --s) {
            const ReconNode nd = nodes[s];
            if (nd.node == root) st[nd.node] = root_state;
            else if (nd.leaf_col >= 0) st[nd.node] = C[(long long)nd.node * Lp];
            else {
                int ps = st[parent[nd.node]];
                if (ps < 0 || ps > B) {
                    err[1] = 1;
                    ps = 0;
                }
                st[nd.node] = C[(long long)nd.node * Lp + ps];
            }
        }
    }
}

hipError_t launch_reconstruct(const ReconNode* nodes, int n_nodes, const int* kids, const int* parent,
                              const double* mt64, int N, int Np, long long slot_elems, const int* counts_T, int F,
                              const int* fam_sorted, int nf, const double* prior, int B, int Rr, double* L_all,
                              int* C_all, double* prod_all, int Lp, int* states, int* err, hipStream_t st)
{
    if (nf <= 0) return hipSuccess;
    if (B + 1 <= 64 * kRecQ && B + 1 <= N) {
        // staged rows: as many as fit in 48 KB, at most 32
        const int jc = std::max(1, std::min(32, (48 * 1024) / (8 * (B + 1))));
        const size_t shm = sizeof(double) * (size_t)jc * (B + 1);
        hipLaunchKernelGGL(k_reconstruct_shared, dim3((nf + kRecW - 1) / kRecW), dim3(64 * kRecW), shm, st, nodes,
                           n_nodes, kids, parent, mt64, N, Np, slot_elems, counts_T, F, fam_sorted, nf, prior, B, Rr,
                           L_all, C_all, prod_all, Lp, states, err, jc);
        return hipGetLastError();
    }
    hipLaunchKernelGGL(k_reconstruct, dim3(nf), dim3(64), 0, st, nodes, n_nodes, kids, parent, mt64, N, Np,
                       slot_elems, counts_T, F, fam_sorted, nf, prior, B, Rr, L_all, C_all, prod_all, Lp, states,
                       err);
    return hipGetLastError();
}

}  // namespace cafe
