// Unit-cube Hex8 element stiffness (E = 1, cell size 1) as a constexpr table.
//
// K = sum over the 2x2x2 Gauss points of B^T D B * detJ (detJ = 1/8, natural-coordinate
// derivatives scaled by 2), with the vertex signs of the eNodMat pattern.  The integrand is a
// polynomial of degree <= 2 per coordinate, so the rule is exact and equals the closed form of
// FEA_VoxelBasedElementStiffnessMatrix (TOP_GPU.m:1465-1503) up to rounding.  Used at run time
// by the host model and at compile time by the finest-level K*u kernels (ν = 0.3), where the
// table's 10 distinct magnitudes become register constants.
#pragma once

#include <hip/hip_runtime.h>

namespace tg {

struct KUnit {
    double v[576];  // column-major 24 x 24
};

__host__ __device__ constexpr KUnit make_kunit(double nu) {
    KUnit K{};
    const double g = 0.57735026918962576451;  // 1/sqrt(3)
    const double sx[8] = {-1, 1, 1, -1, -1, 1, 1, -1};
    const double sy[8] = {-1, -1, 1, 1, -1, -1, 1, 1};
    const double sz[8] = {-1, -1, -1, -1, 1, 1, 1, 1};
    const double c = 1.0 / ((1 + nu) * (1 - 2 * nu));
    double D[6][6] = {};
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) D[i][j] = c * (i == j ? 1 - nu : nu);
    for (int i = 3; i < 6; ++i) D[i][i] = c * (1 - 2 * nu) / 2;
    for (int p = 0; p < 8; ++p) {
        const double xi = sx[p] * g, eta = sy[p] * g, zeta = sz[p] * g;
        double B[6][24] = {};
        for (int a = 0; a < 8; ++a) {
            const double dx = 2 * 0.125 * sx[a] * (1 + sy[a] * eta) * (1 + sz[a] * zeta);
            const double dy = 2 * 0.125 * sy[a] * (1 + sx[a] * xi) * (1 + sz[a] * zeta);
            const double dz = 2 * 0.125 * sz[a] * (1 + sx[a] * xi) * (1 + sy[a] * eta);
            const int iu = 3 * a, iv = iu + 1, iw = iu + 2;
            B[0][iu] = dx;
            B[1][iv] = dy;
            B[2][iw] = dz;
            B[3][iv] = dz; B[3][iw] = dy;
            B[4][iu] = dz; B[4][iw] = dx;
            B[5][iu] = dy; B[5][iv] = dx;
        }
        double DB[6][24] = {};
        for (int i = 0; i < 6; ++i)
            for (int k = 0; k < 24; ++k) {
                double s = 0;
                for (int j = 0; j < 6; ++j) s += D[i][j] * B[j][k];
                DB[i][k] = s;
            }
        for (int col = 0; col < 24; ++col)
            for (int row = 0; row < 24; ++row) {
                double s = 0;
                for (int i = 0; i < 6; ++i) s += B[i][row] * DB[i][col];
                K.v[row + 24 * col] += s * 0.125;
            }
    }
    return K;
}

constexpr KUnit KUNIT_NU03 = make_kunit(0.3);

}  // namespace tg
