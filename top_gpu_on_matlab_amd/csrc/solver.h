// Device solver state (see solver.hip).
#pragma once

#include "coarse.h"
#include "kernels.h"
#include "kfine_tiled.h"
#include "model.h"
#include "pcg.h"

namespace tg {

struct DevLevel {
    Grid g{};
    int span = 0;
    int64_t Nn = 0, Ne = 0, ldk = 0;       // ldk: entry stride of element matrices (padded element count)
    DevBuf<uint8_t> nodeMask, eleMask;
    DevBuf<int32_t> nodFull0, eleFull0;    // solid node / element -> 0-based full-grid index
    DevBuf<double> diag64, r64, x64, t64;  // Jacobi diagonal and V-cycle vectors (full grid, interleaved)
    DevBuf<float> diag32, r32, x32, t32;
    DevBuf<double> Ks64;                   // element matrices, entry-major [576 x ldk]
    DevBuf<double> Ks64p;                  // FP64 mode: their packed upper triangles [300 x ldk] (K*u)
    DevBuf<float> Ks32;                    // mixed mode: the same in FP32
    DevBuf<double> fv64;                   // per-element K*u products, entry-major [24 x ldk]
    DevBuf<float> fv32;
    DevBuf<uint8_t> coeffCnt;              // transferMatCoeffi on the finer grid's nodes (levels >= 2): the
                                           // number of non-empty coarse elements around the node, 1 where unused
    DevBuf<double> P;                      // Pch / Psub [24 x 24 x span^3] (levels >= 2)
};

void build_pch(int span, std::vector<double> &P);

struct Solver {
    const tg_model &model;
    hipStream_t st;
    int L = 0, prec = 0;
    double omega = 0.6;
    bool adapted = false, super = false;
    bool const_k0 = false;  // nu == 0.3: finest K*u uses the compile-time unit element matrix
    int kfine_kernel = 0;   // 0: LDS-tiled element-centric (Walsh basis), 1: node-centric gather
    bool walsh_ok = false;  // the finest element matrix has the isotropic Walsh-basis pattern
    KWVals<double> kw64{};
    KWVals<float> kw32{};
    // optional per-launch timing of the finest-level K*u (HIP events on the solver stream)
    struct TimedLaunch {
        hipEvent_t ev[2] = {nullptr, nullptr};
        int key = 0;  // 4 * (FP32 ? 1 : 0) + epilogue
    };
    int time_fine = 0;  // 1: every tiled finest K*u launch, 2: the plain products (EPI_KU) only
    std::vector<TimedLaunch> tpool;
    size_t tused = 0;
    hipEvent_t *timing_slot(int key);
    int kfine_zchunk = 12;  // node planes per workgroup of the tiled kernel
    int kfine_mfma = 0;     // FP32 tiled finest K*u: 0 Walsh form, 1 / 2 element products on MFMA (FP32 / BF16)
    int kelem_parts = 1;    // threads per element in the FP64 per-element-matrix K*u (1: packed symmetric)
    bool has_modulus = false, assembled = false;
    std::vector<DevLevel> lv;
    CoarseSolver coarse;
    DevBuf<double> K0_64, E64, Esub, Imat, uFree, uFixed;
    DevBuf<float> K0_32, E32;
    DevBuf<int32_t> fullUid;
    DevBuf<double> Gl2, Ech;               // level-2 Galerkin GEMM: G_s (nCh x 576, entry fastest), children moduli
    DevBuf<int64_t> bndList;               // coarse elements with a boundary (fixed-DOF) child
    int64_t nBnd = 0;
    DevBuf<double> b, y;        // PCG right-hand side and iterate (full fine grid)
    PcgWork pw;                 // PCG work vectors
    DevBuf<float> rin32, zout32;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    double last_ms = 0;
    int last_its = 0;

    Solver(const tg_model &m, hipStream_t s);
    ~Solver();
    void set_modulus(const double *dE);
    void assemble();
    template <typename T>
    bool kbyu_level(int i, const T *u, T *out, int epi, const T *rr, const T *dg, int dot_slot = -1,
                    T *xout = nullptr);
    bool fine_tiled() const { return !super && kfine_kernel == 0 && walsh_ok; }
    template <typename T>
    // pk (optional): also write the coarsest level's plane-order copy for its persistent solve
    void restrict_level(int c, const T *rf, T *rc, const PlanePack *pk = nullptr);
    template <typename T>
    void prolong_level(int c, const T *xc, T *xf, bool add);
    template <typename T>
    void vcycle_t(const T *r0, T *z0);
    void finish_dot(const double *parts, int n, int slot);
    // r32_ready: rin32 already holds r0 rounded to FP32 (the CG update wrote it)
    bool vcycle(const double *r0, double *z0, int dot_slot = -1, bool r32_ready = false, bool z32_ok = false,
                bool *z_in32 = nullptr);
    int vc_dot_slot = -1;  // finest post-smoothing also reduces z.r into pw.scal[vc_dot_slot]
    bool vc_dotted = false;
    const double *r64_dot = nullptr;  // mixed: the FP32 finest post-smoothing reduces z.r against this FP64 r
    DevBuf<double> dotp;   // per-workgroup partial sums of the fused dots
    // compliance-gradient work space (tg_compliance_grad)
    DevBuf<double> adjF, adjXp, adjE, adjCe, adjG, adjPart, adjRes;
    void pcg(const double *b_c, double *y_c, double tol, int maxIT, int warm, int *its, double *res);
};

}  // namespace tg

struct tg_solver {
    tg::Solver *s = nullptr;
};
