// Host-side model: voxel discretisation, multigrid hierarchy, boundary conditions.
#pragma once

#include <cstdint>
#include <vector>

#include "tg_internal.h"

namespace tg {

// One level of the hierarchy.  "full grid" arrays cover the level's padded box
// (resY x resX x resZ elements, (resY+1) x (resX+1) x (resZ+1) nodes), column-major (y fastest).
struct HostLevel {
    int32_t resX = 0, resY = 0, resZ = 0, span = 0;
    int64_t Ne = 0, Nn = 0, nInt = 0;
    std::vector<uint8_t> solid;            // full-grid element mask
    std::vector<int32_t> eleMapForward;    // full-grid element -> 1-based solid id (0 void)
    std::vector<int32_t> eleMapBack;       // [Ne] 1-based full-grid element index
    std::vector<int32_t> nodMapForward;    // full-grid node -> 1-based solid node id (0 none)
    std::vector<int32_t> nodFull;          // [Nn] 1-based full-grid node index on THIS level's grid
    std::vector<int32_t> nodMapBack;       // [Nn] as the reference stores it (finest-grid index on levels >= 2)
    std::vector<int32_t> eNodMat;          // [Ne x 8] column-major
    std::vector<int32_t> nodesOnBoundary;  // solid node ids with < 8 adjacent solid elements
    std::vector<int32_t> elementsOnBoundary;  // level 1 only
    // coarse levels (>= 2)
    std::vector<int32_t> upMap;            // [Ne x span^3] column-major
    std::vector<int32_t> transferMat;      // [(span+1)^3 x Ne] column-major
    std::vector<double> coeff;             // transferMatCoeffi [nInt]
    std::vector<int32_t> solidNodeMap;     // solidNodeMapCoarser2Finer [Nn(fine)]
    std::vector<int32_t> fineInterId;      // full fine-grid node -> 1-based intermediate id (0 none)

    int64_t NY() const { return (int64_t)resY + 1; }
    int64_t NX() const { return (int64_t)resX + 1; }
    int64_t NZ() const { return (int64_t)resZ + 1; }
    int64_t nodesFull() const { return NY() * NX() * NZ(); }
    int64_t elemsFull() const { return (int64_t)resY * resX * resZ; }
};

}  // namespace tg

struct tg_model {
    tg_params prm{};
    int32_t nely = 0, nelx = 0, nelz = 0;  // unpadded input dims
    int32_t numLevels = 0, nonDyadic = 0;
    std::vector<tg::HostLevel> lv;
    std::vector<double> Ke;     // [576] column-major, cellSize-scaled
    std::vector<double> Imat;   // [576 x 64] (super-element)
    // conditions
    std::vector<double> fixing;               // [nFix x 4] column-major
    int64_t nFix = 0;
    std::vector<std::vector<double>> loads;   // load cases [nLoad x 4]
    bool finalized = false;
    std::vector<double> F;                    // [3Nn x nCases]
    std::vector<uint8_t> fixedDOF;            // [3Nn]
    std::vector<int32_t> mapUniqueKes;        // [Ne]
    std::vector<int32_t> eleWithFixed;        // [nU]
    std::vector<double> uniqueFree;           // [576 x nU] or [576 x 64 x nU]
    std::vector<double> uniqueFixed;          // [576 x nU]
    std::vector<uint8_t> nodeFixMask;         // [Nn] bit c = DOF c fixed in the element matrices
};

namespace tg {
void element_stiffness(double nu, double cell, double *Ke576);
void super_element_matrices(int num, double cell, double nu, double *out576xq);
void pde_filter_kernel(double radius, double ele_size, double *K64);
// trilinear weight of coarse vertex j at fine position (a,b,c) of a span-s block
double ri_weight(int s, int a, int b, int c, int j);
extern const int VY[8], VX[8], VZ[8];
}  // namespace tg
