// Coarsest-level direct solver (replaces the host CHOLMOD factorisation and the per-V-cycle
// host round trip of TOP_GPU.m:796-802, 1242-1256).
//
// The coarsest grid is cut into node planes along its longest axis, so K_c is block tridiagonal.
// It is factorised from both ends towards a middle separator plane m ("twisted" block Cholesky):
//   chain A, p = 0 .. m-1:     L_pp L_pp^T = A_pp - L_{p,p-1} L_{p,p-1}^T,  L_{p+1,p} = A_{p+1,p} L_pp^-T
//   chain B, p = P-1 .. m+1:   L_pp L_pp^T = A_pp - L_{p,p+1} L_{p,p+1}^T,  L_{p-1,p} = A_{p-1,p} L_pp^-T
//   separator:                 L_mm L_mm^T = A_mm - L_{m,m-1} L_{m,m-1}^T - L_{m,m+1} L_{m,m+1}^T
// on the device (blocked Cholesky: LDS kernels for the diagonal blocks, panels and trailing updates,
// the two chains' planes of a step sharing each launch; rocBLAS syrk/trsm/gemm for the couplings; all on
// the solver stream).  A solve is a batched triangular product y = blockdiag(inv(L_pp)) b, the two
// forward chains advanced together (one launch per step, ~P/2 steps) and the separator, then
// x = blockdiag(inv(L_pp)^T) y and the two backward chains outwards from the separator.  The sequence
// reads only buffers fixed at setup, so it can also be captured once as a HIP graph and replayed.
#pragma once

#include <rocblas/rocblas.h>

#include "kernels.h"
#include "model.h"

namespace tg {

struct CoarseSolver {
    int axis = 0;                   // 0 = y, 1 = x, 2 = z
    int P = 0;                      // number of planes
    std::vector<int64_t> nd;        // DOFs per plane
    std::vector<int64_t> off;       // DOF offset per plane
    std::vector<int64_t> adOff, asOff;  // offsets of the diagonal / sub-diagonal plane blocks
    int m = 0;                          // separator plane; chain A = [0, m), chain B = (m, P-1]
    std::vector<int64_t> fOff, bOff;    // offsets of the padded forward / backward chain operators
    std::vector<int64_t> fLd, bLd;      // their padded row lengths
    int64_t ntot = 0;
    Grid g{};
    DevBuf<int32_t> lidx;           // per full-grid node: local node index within its plane (-1 none)
    DevBuf<int64_t> d_off, d_adOff, d_asOff;
    int64_t npairs = 0;             // row pairs (r, n_p - 1 - r) over all planes
    DevBuf<int32_t> pairPlane;      // plane of each row pair
    DevBuf<int64_t> d_pairOff;      // first row pair of each plane
    DevBuf<double> Ad, As, Linv, LT, Fop, Bop;
    DevBuf<double> bb, yy, xx;
    DevBuf<rocblas_int> info;
    DevBuf<double> chW;             // inv(L_kk) of the current Cholesky panel
    rocblas_handle h = nullptr;
    bool factored = false;
    hipGraphExec_t exec = nullptr;  // the captured solve sequence
    bool use_graph = false;         // replay the captured sequence (false: launch it directly)

    ~CoarseSolver();
    void setup(const HostLevel &L, const Grid &grid, hipStream_t st);
    // Assemble the plane blocks from the level's element matrices (entry-major, ldk) and factorise.
    void factor(const double *Ks, int64_t ldk, const uint8_t *eleMask, const uint8_t *nodeMask, hipStream_t st);
    // x = K_c^{-1} r on full-grid interleaved vectors (T = float or double; solve in FP64).
    template <typename T>
    void solve(const T *r, T *x, const uint8_t *nodeMask, hipStream_t st);
    void enqueue_core(hipStream_t s) const;
};

}  // namespace tg
