// Device objects of the optimisation layer (see topopt.hip).
#pragma once

#include "kernels.h"
#include "model.h"
#include "pcg.h"

namespace tg {

// Nodal Helmholtz (PDE) density filter on the finest grid.
struct Filter {
    const tg_model &m;
    hipStream_t st;
    Grid g{};
    int64_t Ne = 0;
    DevBuf<double> Kd, inv, b, y;
    DevBuf<uint8_t> nodeMask, eleMask;
    DevBuf<int32_t> eleFwd, eleFull0;
    PcgWork pw;
    Filter(const tg_model &mm, double rMin, hipStream_t s);
    void apply(const double *src, double *tar, double tol, int maxIT, int *its, double *res);
};

// imfilter density filter of super-element mode on the num-times refined voxel grid.
struct ConvFilter {
    int num, X = 0, Y = 0, Z = 0, GY = 0, GX = 0, GZ = 0, hr = 0;
    hipStream_t st;
    DevBuf<double> hd, Hs, tmp;
    ConvFilter(int nelx, int nely, int nelz, double r, int n, hipStream_t s);
    void forward(const double *x, double *out);
    void adjoint(const double *dc, double *out);
};

}  // namespace tg

struct tg_filter {
    tg::Filter *f = nullptr;
};

struct tg_convfilter {
    tg::ConvFilter *f = nullptr;
};
