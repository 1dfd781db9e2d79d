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

hipError_t launch_reconstruct(const ReconNode* nodes, int n_nodes, const int* kids, const int* parent,
                              const double* mt64, int N, int Np, long long slot_elems, const int* counts_T, int F,
                              const int* fam_sorted, int nf, const double* prior, int B, int Rr, double* L_all,
                              int* C_all, double* prod_all, int Lp, int* states, int* err, hipStream_t st)
{
    if (nf <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_reconstruct, dim3(nf), dim3(64), 0, st, nodes, n_nodes, kids, parent, mt64, N, Np,
                       slot_elems, counts_T, F, fam_sorted, nf, prior, B, Rr, L_all, C_all, prod_all, Lp, states,
                       err);
    return hipGetLastError();
}

}  // namespace cafe
