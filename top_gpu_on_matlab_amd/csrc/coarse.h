// Coarsest-level direct solver (replaces the host CHOLMOD factorisation and the per-V-cycle
// host round trip of TOP_GPU.m:796-802, 1242-1256).
//
// The coarsest grid is cut into node planes along its longest axis, so K_c is block tridiagonal with
// dense diagonal blocks A_pp and sparse couplings A_{p,p+-1} (27 entries per row: the 3 x 3 node
// neighbourhood in the next plane, 3 DOFs each).  It is eliminated from both ends towards a middle
// separator plane m ("twisted" block LDL'), keeping the inverse of every plane's Schur complement:
//   chain A, p = 0 .. m-1:    S_p = A_pp - A_{p,p-1} N_{p-1},     N_p = S_p^-1 A_{p,p+1}
//   chain B, q = P-1 .. m+1:  S_q = A_qq - A_{q,q+1} N_{q+1},     N_q = S_q^-1 A_{q,q-1}
//   separator:                S_m = A_mm - A_{m,m-1} N_{m-1} - A_{m,m+1} N_{m+1}
// Each S_p^-1 is formed in place by a blocked Gauss-Jordan sweep (64-wide pivot blocks: the pivot
// block's inverse in LDS, the column panel, then a rank-64 update of the whole plane block on the
// matrix cores); the two chains' planes of a step share every launch.  The solve operators are
//   forward  F_p = S_p^-1 A_{p,p-1} (chain A), S_q^-1 A_{q,q+1} (chain B), [S_m^-1 A_{m,m-1} | S_m^-1 A_{m,m+1}]
//   backward B_p = N_p
// and a solve is c = blockdiag(S_p^-1) b, the forward recurrences z_p = c_p - F_p z_{p-1|p+1} of both
// chains advanced together, the separator x_m = c_m - F_m [z_{m-1}; z_{m+1}], then
// x_p = z_p - B_p x_{p+1|p-1} outwards.  By default the whole sequence is one persistent launch whose
// workgroups pass the recurrence vectors to each other (k_coarse_persistent, in coarse.hip); step by
// step, it is one launch for c and one per step, and that sequence reads only buffers fixed at setup,
// so it can also be captured once as a HIP graph and replayed.  Everything is FP64 and deterministic
// (fixed summation orders, no floating-point atomics).
#pragma once

#include "kernels.h"
#include "model.h"

namespace tg {

// "Not written yet" in the persistent solve's recurrence vectors: a signalling NaN, which arithmetic
// never produces.
constexpr unsigned long long COARSE_EMPTY = 0x7ff4c0a5e5e17e11ull;

struct CoarseSolver {
    int axis = 0;                   // 0 = y, 1 = x, 2 = z
    int P = 0;                      // number of planes
    std::vector<int64_t> nd;        // DOFs per plane
    std::vector<int64_t> npad;      // DOFs per plane padded to a multiple of 64 (the sweep's blocks)
    std::vector<int64_t> off;       // DOF offset per plane
    std::vector<int64_t> sOff;      // offset of each plane's npad x npad block (row-major)
    int m = 0;                          // separator plane; chain A = [0, m), chain B = (m, P-1]
    std::vector<int64_t> fOff, bOff;    // offsets of the padded forward / backward operators
    std::vector<int64_t> fLd, bLd;      // their padded row lengths
    int64_t ntot = 0;
    Grid g{};
    DevBuf<int32_t> lidx;           // per full-grid node: local node index within its plane (-1 none)
    DevBuf<int32_t> gmap;           // per plane-order DOF: its full-grid DOF
    DevBuf<int32_t> gzero;          // full-grid DOFs of inactive nodes (zero in every solution)
    int64_t nzero = 0;
    DevBuf<int64_t> d_off, d_sOff, d_npad, d_nd;
    DevBuf<int32_t> rowPlane;       // plane of each DOF row (solve-time block products)
    DevBuf<double> Sinv;            // per plane: A_pp, then S_p, then S_p^-1 (row-major, padded)
    DevBuf<double> dnV, upV;        // couplings to the previous / next plane, 27 per DOF row
    DevBuf<int32_t> dnC, upC;       // their column (local DOF in that plane)
    DevBuf<double> Fop, Bop;
    DevBuf<double> T, Rw, Dk;       // sweep work: column panel, old pivot row, pivot inverse (2 planes)
    DevBuf<double> bb, yy;
    DevBuf<double> xx;  // the persistent solve's separator and backward results (yy: its forward ones)
    DevBuf<int> info;
    bool factored = false;
    hipGraphExec_t exec = nullptr;  // the captured solve sequence
    bool use_graph = false;         // replay the captured sequence (false: launch it directly)
    // the solve as one persistent launch (stages passing the recurrence vectors directly); the step-by-step
    // sequence above is used when this is off, when a graph is requested, or when the planes are too
    // wide for it (more than 1024 DOFs per plane)
    bool use_persistent = true;
    bool pers_streaming = true;  // non-temporal loads of the factor rows
    unsigned pers_polls = 1u << 17;  // polls of an input (~1 us apart) before a stage gives up (~0.1 s)
    bool pers_built = false, pers_ok = false;
    int pers_nk = 0, pers_groups = 0, pers_stages = 0, pers_L1 = 0, pers_L2 = 0;
    DevBuf<unsigned char> pers_table;        // the stage table (device copy)
    DevBuf<int> pers_err;                    // set when a stage gave up waiting for its input
    bool pers_pending = false;               // launched since the last flag check

    ~CoarseSolver();
    void setup(const HostLevel &L, const Grid &grid, hipStream_t st);
    // Assemble the plane blocks from the level's element matrices (entry-major, ldk) and factorise.
    void factor(const double *Ks, int64_t ldk, const uint8_t *eleMask, const uint8_t *nodeMask, hipStream_t st);
    // x = K_c^{-1} r on full-grid interleaved vectors (T = float or double; solve in FP64).  packed: the
    // plane-order right-hand side and the sentinel fill were written by the kernel that formed r (as
    // plane_pack below describes), so the solve does not pack them again.
    template <typename T>
    void solve(const T *r, T *x, const uint8_t *nodeMask, hipStream_t st, bool packed = false);
    // Whether the next solve runs as the persistent launch; if so, pk tells the kernel that forms the
    // right-hand side where to put its plane-order copy.
    bool plane_pack(PlanePack &pk, hipStream_t st);
    void enqueue_core(hipStream_t s) const;
    void build_persistent(hipStream_t st);
    void enqueue_persistent(hipStream_t st, double *x64, float *x32);
    // Fails (TG_ERR_HIP) if a persistent launch since the last check gave up waiting for an input (its
    // result is then invalid); synchronises the stream when there is something to check.
    void check_barrier(hipStream_t st);
};

}  // namespace tg
