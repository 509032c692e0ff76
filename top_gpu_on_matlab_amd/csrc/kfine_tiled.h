// Finest-level K*u for standard voxels: element-centric, LDS-tiled z-march, with the element
// matrix applied in the Walsh (+-1 Hadamard) basis of the 8 vertices.
//
// Walsh factorisation.  Number the vertices q = vy + 2 vx + 4 vz and let W[q][m] = prod over the
// axes a in m of (2 q_a - 1) (the 8 modes: constant, the three linear, three bilinear and the
// trilinear function, evaluated at the vertices).  W^T W = 8 I, so for the 24x24 element matrix
//     K0 = T KW T^T,   T = W (x) I_3,   KW = T^T K0 T / 64.
// For any isotropic Hex8 matrix KW has 45 nonzeros of 576 (the constant mode carries no energy and
// the cube's symmetries decouple most mode pairs), so f_e = E_e K0 u_e costs an 8-point butterfly per
// component (72 adds), 45 FMAs, 21 scalings and the inverse butterflies (below: 24 adds per element
// for the z stage and 24 per face for the y and x stages) instead of 576 FMAs.
//
// Tiling.  A workgroup owns a KT_TY x KT_TX column of nodes (y, x) over a chunk of z node planes
// and marches through the element layers between them.  Per element layer:
//   1. the next node plane of the (KT_TY+2) x (KT_TX+2) halo box (fetched into registers during
//      the previous layer) is committed to LDS, fixed DOFs zeroed (the zero columns of the
//      BC-modified element matrices, TOP_GPU.m:704-720), and the plane after it is fetched;
//   2. each of the 256 lanes computes the element vector of one element of the (KT_TY+1) x
//      (KT_TX+1) layer in the Walsh basis and splits it by the inverse z stage into the values of its
//      lower and upper faces (face Walsh basis).  The lower face is the face of the same (y, x) that
//      the lane's element of the previous layer left as its upper face: the two are added, the sum is
//      taken through the inverse y and x stages to its 4 vertex values, which go to LDS; the upper
//      face is kept in registers for the next layer.  Each face is thus inverted once, not once per
//      element on either side of it;
//   3. each owned node sums the vertex values of the 4 faces around it (fixed order: deterministic),
//      applies identity rows for fixed DOFs and the epilogue (K*u, r - K*u or the damped-Jacobi
//      update), and writes its 3 values once.
#pragma once

#include "kernels.h"

namespace tg {

constexpr int KT_TY = 15, KT_TX = 15;                // owned nodes per plane (225 of 256 lanes)
constexpr int KT_PY = KT_TY + 2, KT_PX = KT_TX + 2;  // staged nodes per plane, with the halo
constexpr int KT_EY = KT_TY + 1, KT_EX = KT_TX + 1;  // elements per layer: 256, one per lane
constexpr int KT_PN = KT_PY * KT_PX;
constexpr int KT_NI = (3 * KT_PN + 255) / 256;        // staged values per lane and plane
static_assert(KT_EY * KT_EX == 256, "one element per lane");

// ---- Walsh basis ---------------------------------------------------------------------------
__host__ __device__ constexpr int vtx_q(int b) { return vtx_y(b) + 2 * vtx_x(b) + 4 * vtx_z(b); }
__host__ __device__ constexpr int walsh_sign(int q, int m) {
    int s = 1;
    for (int a = 0; a < 3; ++a)
        if ((m >> a) & 1) s *= ((q >> a) & 1) ? 1 : -1;
    return s;
}

// KW = T^T K0 T / 64 (column-major over index 3m + c), entries below 1e-13 max|KW| set to zero.
// Two passes of 8-term sums (A = K0 T, then T^T A) keep the compile-time evaluation short.
inline constexpr KUnit make_kwalsh(const KUnit &K) {
    int sg[8][8] = {};  // sg[b][m] = W[q(b)][m]
    for (int b = 0; b < 8; ++b)
        for (int m = 0; m < 8; ++m) sg[b][m] = walsh_sign(vtx_q(b), m);
    KUnit A{}, w{};
    for (int j = 0; j < 24; ++j)
        for (int i = 0; i < 24; ++i) {
            double s = 0;
            for (int b = 0; b < 8; ++b) s += K.v[i + 24 * (3 * b + j % 3)] * sg[b][j / 3];
            A.v[i + 24 * j] = s;
        }
    double mx = 0;
    for (int j = 0; j < 24; ++j)
        for (int i = 0; i < 24; ++i) {
            double s = 0;
            for (int b = 0; b < 8; ++b) s += sg[b][i / 3] * A.v[(3 * b + i % 3) + 24 * j];
            s /= 64;
            w.v[i + 24 * j] = s;
            mx = s > mx ? s : (-s > mx ? -s : mx);
        }
    for (int k = 0; k < 576; ++k)
        if (w.v[k] < 1e-13 * mx && w.v[k] > -1e-13 * mx) w.v[k] = 0;
    return w;
}

struct KWPattern {
    int idx[576];  // compact position of entry (i, j) = i + 24 j, or -1
    int n;
};
inline constexpr KWPattern make_kwpattern(const KUnit &kw) {
    KWPattern p{};
    p.n = 0;
    for (int k = 0; k < 576; ++k) p.idx[k] = kw.v[k] != 0 ? p.n++ : -1;
    return p;
}
constexpr KUnit KWALSH_NU03 = make_kwalsh(KUNIT_NU03);
__host__ __device__ constexpr KWPattern KW_PATTERN = make_kwpattern(KWALSH_NU03);
constexpr int KW_NNZ = 45;
static_assert(KW_PATTERN.n == KW_NNZ, "isotropic Hex8 matrix: 45 nonzeros in the Walsh basis");

template <typename T>
struct KWVals {
    T v[48];  // the KW_NNZ pattern entries of cs * KW, in pattern order
};

// ---- tiles -----------------------------------------------------------------------------------
struct KTiles {
    int tilesY, tilesX, chunks, zchunk;
    __host__ __device__ int64_t count() const { return (int64_t)tilesY * tilesX * chunks; }
};

inline KTiles ktiles_for(const Grid &g, int zchunk) {
    KTiles t;
    t.tilesY = (g.NY + KT_TY - 1) / KT_TY;
    t.tilesX = (g.NX + KT_TX - 1) / KT_TX;
    t.zchunk = zchunk;
    t.chunks = (g.NZ + zchunk - 1) / zchunk;
    return t;
}

// 8-point Walsh butterflies over q: forward m = W^T x, inverse x = W g.
template <typename T>
__device__ __forceinline__ void walsh_fwd(T x[8]) {
#pragma unroll
    for (int s = 1; s < 8; s <<= 1)
#pragma unroll
        for (int i = 0; i < 8; ++i)
            if (!(i & s)) {
                const T a = x[i], b = x[i | s];
                x[i] = a + b;
                x[i | s] = b - a;
            }
}
template <typename T>
__device__ __forceinline__ void walsh_inv(T x[8]) {
#pragma unroll
    for (int s = 1; s < 8; s <<= 1)
#pragma unroll
        for (int i = 0; i < 8; ++i)
            if (!(i & s)) {
                const T a = x[i], b = x[i | s];
                x[i] = a - b;
                x[i | s] = a + b;
            }
}

// CK: KW is the compile-time unit matrix for nu = 0.3 and each element is scaled by cs * E_e;
// otherwise KW (cellSize included) comes from kw and cs is ignored.
// DOT: also write this workgroup's share of a CG dot product to partial[tile] (FP64, fixed order):
// u.out for the plain product (p'Kp), r.out for the Jacobi epilogue (z'r).
// EPI_RESID_JAC0: the V-cycle's pre-smoothing fused in: the operand is x = omega r ./ diag (formed
// while staging; u is not read), x is written to xout and out = r - K x.
// R64 (FP32 damped Jacobi with DOT only): the dot is taken against rdot, the FP64 residual, instead of
// the FP32 one -- the mixed V-cycle's last step, giving the FP64 CG its z.r (z itself stays FP32: its
// widening to FP64 is exact and happens where the CG reads it).
// MF (FP32 only; 0 = the Walsh form): the element product runs on the matrix cores as a dense 24 x 24
// product (mfma_K0: the element matrix, cellSize included, column-major FP32; kw and cs are unused).
// MF = 1: MFMA f32 32x32x2, FP32 operands.  MF = 2: MFMA f32 32x32x16 bf16, operands (E_e u_e and K0)
// rounded to BF16, products accumulated in FP32.  MF = 3: as 2 with each operand split into two BF16
// parts (a = hi + lo, lo = bf16(a - hi), 16 significant bits together) and three products per block,
// hi lo + lo hi + hi hi (lo lo, below 2^-16 of the product, is dropped).  Wave w takes the layer's elements 64 w .. 64 w + 63 as
// two blocks of 32 (columns of the product).  Lane (n, h) (n = lane % 32, h = lane / 32) supplies the 12
// values of face h (h = 0: node plane ez, 1: ez + 1) of element n, read from LDS: with MF = 1 as input
// DOFs 2 s + h (value s = vertex s / 3 of the face, component s % 3), 12 steps per block; with MF = 2 as
// input DOFs 16 s + 8 h + j (value 8 s + j, j < 8; values past 11 are zero padding), 2 steps per block.
// Output rows are permuted so that lane half h holds, in result registers 0..11, the 12 values of face
// h ^ (ez & 1) of its element; the permutation alternates with the layer parity, so the upper face a
// lane computes in one layer and the lower face of the element above it in the next layer fall to the
// same lane and are summed in registers, as the Walsh form does.  Lane layouts: lane l holds A(l % 32,
// K-block l / 32) and B(K-block l / 32, l % 32) (one value per lane for f32 32x32x2, eight consecutive K
// for bf16 32x32x16); result register 4 g + i is at row 8 g + 4 (l / 32) + i, column l % 32.
// Measured at 128 x 256 x 128 the dense products are slower than the Walsh form (MF = 1: 141 / 174 us
// against 94 / 89 us for the residual / smoothing applies): 1536 padded flops per element on the FP32
// matrix pipe take >= 58 us alone, where the Walsh form spends ~250 VALU operations per element.
template <typename T, int EPI, bool CK, bool DOT, bool R64 = false, int MF = 0>
__global__ __launch_bounds__(256, MF == 3 ? 2 : ((sizeof(T) == 8 || MF != 0) ? 3 : 4)) void k_kfine_walsh(const T *__restrict__ u, T *__restrict__ out,
                                                     const T *__restrict__ E, const uint8_t *__restrict__ mask,
                                                     const KWVals<T> kw, T cs, const Grid g, const KTiles kt,
                                                     const T *__restrict__ r, const T *__restrict__ diag, T omega,
                                                     double *__restrict__ partial, T *__restrict__ xout,
                                                     const double *__restrict__ rdot,
                                                     const float *__restrict__ mfma_K0 = nullptr) {
    static_assert(!R64 || (EPI == EPI_JACOBI && DOT && sizeof(T) == 4), "R64: FP32 Jacobi with the dot only");
    static_assert(MF == 0 || sizeof(T) == 4, "MF: FP32 only");
    constexpr bool J0 = EPI == EPI_RESID_JAC0;
    __shared__ T sU[2][3][KT_PN];
    __shared__ uint8_t sM[2][KT_PN];
    // face values: the 4 vertex contributions x 3 components of each lane's face, [3 q + c][lane]
    __shared__ T sF[12][256];
    const int t = threadIdx.x;
    // XCD-aware tile order: each XCD takes a contiguous range of tiles (xcd_contiguous)
    int bid = xcd_contiguous((int)blockIdx.x, (int)gridDim.x);
    const int tile = bid;  // the DOT partial goes to partial[tile]: the sum's order follows the tiles
    const int tpc = kt.tilesY * kt.tilesX;
    const int tz = bid / tpc;
    bid -= tz * tpc;
    const int ty = bid % kt.tilesY, tx = bid / kt.tilesY;
    const int y0 = ty * KT_TY, x0 = tx * KT_TX;
    const int z0 = tz * kt.zchunk;
    const int z1 = min(z0 + kt.zchunk, g.NZ);
    const int64_t NYX = (int64_t)g.NY * g.NX;
    const int64_t EYX = (int64_t)g.EY * g.EX;

    // Addressing: every load is unconditional, at a 32-bit lane offset inside a node (element)
    // plane from a plane base that is uniform across the workgroup; coordinates outside the grid
    // are clamped and the loaded value is masked when it is committed.  Straight-line loads keep
    // the memory-counter waits static: a wait for one group of loads never drains a younger one.
    // element lane
    const int ly = t % KT_EY, lx = t / KT_EY;
    const int ey = y0 - 1 + ly, ex = x0 - 1 + lx;
    const bool eIn = (unsigned)ey < (unsigned)g.EY && (unsigned)ex < (unsigned)g.EX;
    const int offE = min(max(ey, 0), g.EY - 1) + g.EY * min(max(ex, 0), g.EX - 1);
    // node lane (lanes >= KT_TY * KT_TX own none)
    const int oy = t % KT_TY, ox = t / KT_TY;
    const bool owner = t < KT_TY * KT_TX && y0 + oy < g.NY && x0 + ox < g.NX;
    const int offO = 3 * (min(y0 + oy, g.NY - 1) + g.NY * min(x0 + ox, g.NX - 1));
    // staged values of the halo plane: value i = t + 256 k is component i % 3 of node i / 3
    int offU[KT_NI];
    bool okXY[KT_NI];
#pragma unroll
    for (int k = 0; k < KT_NI; ++k) {
        const int i = min(t + 256 * k, 3 * KT_PN - 1);
        const int node = i / 3, c = i - 3 * node;
        const int gy = y0 - 1 + node % KT_PY, gx = x0 - 1 + node / KT_PY;
        okXY[k] = (unsigned)gy < (unsigned)g.NY && (unsigned)gx < (unsigned)g.NX;
        offU[k] = 3 * (min(max(gy, 0), g.NY - 1) + g.NY * min(max(gx, 0), g.NX - 1)) + c;
    }
    auto zclamp = [&](int z, int n) { return min(max(z, 0), n - 1); };

    // next plane's staged values (registers until committed)
    T pre[KT_NI], preD[J0 ? KT_NI : 1];
    uint8_t preM[KT_NI];
    bool preZ = false;
    auto fetch = [&](int z) {
        const int64_t b = NYX * zclamp(z, g.NZ);
        const T *src = (J0 ? r : u) + 3 * b;
#pragma unroll
        for (int k = 0; k < KT_NI; ++k) {
            pre[k] = src[offU[k]];
            if (J0) preD[J0 ? k : 0] = diag[3 * b + offU[k]];
            preM[k] = mask[b + offU[k] / 3];
        }
        preZ = (unsigned)z < (unsigned)g.NZ;
    };
    auto commit = [&](int slot) {
#pragma unroll
        for (int k = 0; k < KT_NI; ++k) {
            const int i = t + 256 * k;
            if (i < 3 * KT_PN) {
                const int node = i / 3, c = i - 3 * node;
                const bool ok = okXY[k] && preZ;
                const uint8_t mk = ok ? preM[k] : (uint8_t)0;
                const T v = J0 ? omega * (pre[k] / preD[J0 ? k : 0]) : pre[k];
                sU[slot][c][node] = (ok && !(mk & (1 << c))) ? v : (T)0;
                if (c == 0) sM[slot][node] = mk;
            }
        }
    };
    T Eraw = 0;
    bool Eok = false;
    auto fetchE = [&](int layer) {
        Eraw = E[EYX * zclamp(layer, g.EZ) + offE];
        Eok = eIn && (unsigned)layer < (unsigned)g.EZ;
    };
    // the owned node's own values for the epilogue (raw u: identity rows and the dot; r; diag),
    // loaded one plane ahead, after this plane's stores.  u is read from memory again rather than from
    // the LDS stage: the stage has fixed DOFs zeroed, and patching them in costs the FP64 operator more
    // (a select on a fresh load waits for it, draining the next plane's prefetch; measured 104 against
    // 95 us at 128 x 256 x 128) than the 12 or 24 bytes per node the reload reads.
    constexpr bool NEED_U = EPI == EPI_KU || EPI == EPI_RESID || EPI == EPI_JACOBI;
    constexpr bool NEED_R = EPI != EPI_KU;
    constexpr bool NEED_D = EPI == EPI_JACOBI || J0;
    T oU[3] = {0, 0, 0}, oR[3] = {0, 0, 0}, oD[3] = {1, 1, 1};
    double oW[3] = {0, 0, 0};  // R64: the FP64 residual for the dot
    auto fetchO = [&](int z) {
        const int64_t b = 3 * NYX * zclamp(z, g.NZ) + offO;
#pragma unroll
        for (int d = 0; d < 3; ++d) {
            if (NEED_U) oU[d] = u[b + d];
            if (NEED_R) oR[d] = r[b + d];
            if (NEED_D) oD[d] = diag[b + d];
            if (R64) oW[d] = rdot[b + d];
        }
    };

    fetch(z0 - 1);
    commit(0);
    fetch(z0);
    fetchE(z0 - 1);
    fetchO(z0 - 1);
    // the upper face of this lane's element (the face of the same (y, x) in the next node plane), in
    // the face's Walsh basis [3 i + c]: completed by the next layer's element before it is inverted
    T kup[12] = {0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0};
    // FP32 (register room for it): the forward y and x butterflies of the element's lower face are
    // those its upper face had in the previous layer, kept in klo [3 q + c]; each face is transformed
    // once.  FP64 transforms both faces of every element (12 more registers would cost a wave).
    constexpr bool KEEPF = sizeof(T) == 4 && MF == 0;
    T klo[KEEPF ? 12 : 1];
    // MF: the lane's A operand for both layer parities, and the face kept from the previous layer
    typedef __bf16 mf_bf8 __attribute__((ext_vector_type(8)));
    const int ml = t & 63, mn = ml & 31, mh = ml >> 5, mw = t >> 6;
    float mA[MF == 1 ? 2 : 1][MF == 1 ? 12 : 1];
    mf_bf8 mAb[MF >= 2 ? 2 : 1][MF >= 2 ? 2 : 1];
    mf_bf8 mAl[MF == 3 ? 2 : 1][MF == 3 ? 2 : 1];  // MF = 3: the low BF16 parts (a = hi + lo)
    float mkept[MF ? 2 : 1][MF ? 12 : 1];
    if constexpr (MF != 0) {
        const int g8 = mn >> 3, hh = (mn >> 2) & 1, ii = mn & 3, j = 4 * g8 + ii;  // output row mn
        // entry of the permuted K0: output row mn at parity P, input value v (0..11) of face mh
        auto aval = [&](int P, int v) -> float {
            if (j >= 12 || v >= 12) return 0.0f;
            const int qi = v / 3, ci = v % 3, qo = j / 3, co = j % 3;
            const int col = 3 * vtx_of(qi & 1, qi >> 1, mh) + ci;
            const int row = 3 * vtx_of(qo & 1, qo >> 1, hh ^ P) + co;
            return mfma_K0[row + 24 * col];
        };
#pragma unroll
        for (int P = 0; P < 2; ++P) {
            if constexpr (MF == 1) {
#pragma unroll
                for (int s2 = 0; s2 < 12; ++s2) mA[P][s2] = aval(P, s2);
            } else {
#pragma unroll
                for (int s2 = 0; s2 < 2; ++s2)
#pragma unroll
                    for (int jj = 0; jj < 8; ++jj) {
                        const float a = aval(P, 8 * s2 + jj);
                        mAb[P][s2][jj] = (__bf16)a;
                        if constexpr (MF == 3) mAl[P][s2][jj] = (__bf16)(a - (float)mAb[P][s2][jj]);
                    }
            }
        }
#pragma unroll
        for (int b = 0; b < 2; ++b)
#pragma unroll
            for (int s2 = 0; s2 < 12; ++s2) mkept[b][s2] = 0.0f;
    }
    auto face_fwd = [&](int slot, T *f) {  // f[3 q + c] = face (ly, lx) of a staged plane, y and x stages
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            T x[4];
#pragma unroll
            for (int q = 0; q < 4; ++q) x[q] = sU[slot][c][(ly + (q & 1)) + KT_PY * (lx + (q >> 1))];
#pragma unroll
            for (int st = 1; st < 4; st <<= 1)
#pragma unroll
                for (int q = 0; q < 4; ++q)
                    if (!(q & st)) {
                        const T a = x[q], b = x[q | st];
                        x[q] = a + b;
                        x[q | st] = b - a;
                    }
#pragma unroll
            for (int q = 0; q < 4; ++q) f[3 * q + c] = x[q];
        }
    };
    double dsum = 0;        // DOT: this lane's share, accumulated plane by plane
    int cur = 0;
    for (int ez = z0 - 1; ez < z1; ++ez) {
        commit(cur ^ 1);  // node plane ez + 1
        const T Ee = Eok ? Eraw : (T)0;
        lds_barrier();
        // the owned node's mask (slot cur is overwritten once every lane has passed the next barrier)
        const uint8_t mk = owner ? sM[cur][(oy + 1) + KT_PY * (ox + 1)] : (uint8_t)0;
        fetch(ez + 2);  // past the chunk on the last layer: clamped, never committed
        fetchE(ez + 1);
        // ---- element vector of this lane's element, passed on as face values
        if constexpr (MF != 0) {
            typedef float mf_f16 __attribute__((ext_vector_type(16)));
            const int P = ez & 1;
            // E of this lane's two elements: its own (block mh) and its partner lane's (block 1 - mh)
            const float Eo = (float)Ee, Ep = __shfl_xor(Eo, 32, 64);
            mf_f16 acc[2];
#pragma unroll
            for (int b = 0; b < 2; ++b) {
                const int te = 64 * mw + 32 * b + mn;
                const int eyb = te % KT_EY, exb = te / KT_EY;
                const float Eb = b == mh ? Eo : Ep;
                float bv[12];
#pragma unroll
                for (int s2 = 0; s2 < 12; ++s2) {
                    const int qi = s2 / 3, ci = s2 % 3;
                    bv[s2] = Eb * (float)sU[mh ? cur ^ 1 : cur][ci][(eyb + (qi & 1)) + KT_PY * (exb + (qi >> 1))];
                }
#pragma unroll
                for (int k = 0; k < 16; ++k) acc[b][k] = 0.0f;
                if constexpr (MF == 1) {
#pragma unroll
                    for (int s2 = 0; s2 < 12; ++s2)
                        acc[b] = __builtin_amdgcn_mfma_f32_32x32x2f32(P ? mA[1][s2] : mA[0][s2], bv[s2], acc[b], 0, 0, 0);
                } else {
                    mf_bf8 b0, b1;
#pragma unroll
                    for (int jj = 0; jj < 8; ++jj) {
                        b0[jj] = (__bf16)bv[jj];
                        b1[jj] = (__bf16)(jj < 4 ? bv[8 + jj] : 0.0f);
                    }
                    if constexpr (MF == 3) {  // the cross terms first: hi lo + lo hi, then hi hi
                        mf_bf8 l0, l1;
#pragma unroll
                        for (int jj = 0; jj < 8; ++jj) {
                            l0[jj] = (__bf16)(bv[jj] - (float)b0[jj]);
                            l1[jj] = (__bf16)(jj < 4 ? bv[8 + jj] - (float)b1[jj] : 0.0f);
                        }
                        acc[b] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(P ? mAl[1][0] : mAl[0][0], b0, acc[b], 0, 0, 0);
                        acc[b] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(P ? mAl[1][1] : mAl[0][1], b1, acc[b], 0, 0, 0);
                        acc[b] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(P ? mAb[1][0] : mAb[0][0], l0, acc[b], 0, 0, 0);
                        acc[b] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(P ? mAb[1][1] : mAb[0][1], l1, acc[b], 0, 0, 0);
                    }
                    acc[b] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(P ? mAb[1][0] : mAb[0][0], b0, acc[b], 0, 0, 0);
                    acc[b] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(P ? mAb[1][1] : mAb[0][1], b1, acc[b], 0, 0, 0);
                }
            }
            // lanes holding the lower face (mh == P) add the face kept from the layer below and pass it on
#pragma unroll
            for (int b = 0; b < 2; ++b) {
                const int te = 64 * mw + 32 * b + mn;
#pragma unroll
                for (int k = 0; k < 12; ++k) {
                    const float v = acc[b][k] + mkept[b][k];
                    if (mh == P) sF[k][te] = (T)v;
                    mkept[b][k] = acc[b][k];
                }
            }
        } else {
            const T Es = CK ? cs * Ee : Ee;
            T m[24];
            if (KEEPF) {
                if (ez == z0 - 1) face_fwd(cur, klo);
                T up[12];
                face_fwd(cur ^ 1, up);
                // forward z stage (as in walsh_fwd); the upper face is the next layer's lower face
#pragma unroll
                for (int i = 0; i < 12; ++i) {
                    const T a = klo[i], b = up[i];
                    m[i] = a + b;
                    m[12 + i] = b - a;
                    klo[i] = b;
                }
            } else {
#pragma unroll
                for (int c = 0; c < 3; ++c) {
                    T x[8];
#pragma unroll
                    for (int b = 0; b < 8; ++b) {
                        const int pn = (ly + vtx_y(b)) + KT_PY * (lx + vtx_x(b));
                        x[vtx_q(b)] = sU[vtx_z(b) ? cur ^ 1 : cur][c][pn];
                    }
                    walsh_fwd(x);
#pragma unroll
                    for (int q = 0; q < 8; ++q) m[3 * q + c] = x[q];
                }
            }
            T gq[24];
#pragma unroll
            for (int i = 0; i < 24; ++i) {
                T acc = 0;
#pragma unroll
                for (int j = 0; j < 24; ++j)
                    if (KW_PATTERN.idx[i + 24 * j] >= 0)
                        acc += (CK ? (T)KWALSH_NU03.v[i + 24 * j] : kw.v[KW_PATTERN.idx[i + 24 * j]]) * m[j];
                gq[i] = acc * Es;
            }
            // inverse z stage: modes q and q + 4 give the lower- and upper-face values of face mode q;
            // the lower face (node plane ez) adds to what the element below left for it
#pragma unroll
            for (int c = 0; c < 3; ++c) {
                T x[4];
#pragma unroll
                for (int q = 0; q < 4; ++q) {
                    const T a = gq[3 * q + c], b = gq[3 * (q + 4) + c];
                    x[q] = kup[3 * q + c] + (a - b);
                    kup[3 * q + c] = a + b;
                }
                // inverse y and x stages of the face: vertex values, vertex q = vy + 2 vx
#pragma unroll
                for (int st = 1; st < 4; st <<= 1)
#pragma unroll
                    for (int q = 0; q < 4; ++q)
                        if (!(q & st)) {
                            const T a = x[q], b = x[q | st];
                            x[q] = a - b;
                            x[q | st] = a + b;
                        }
#pragma unroll
                for (int q = 0; q < 4; ++q) sF[3 * q + c][t] = x[q];
            }
        }
        lds_barrier();
        // ---- owned node: complete plane ez from the 4 faces around it
        if (owner) {
            T lo[3] = {0, 0, 0};
#pragma unroll
            for (int dx = 0; dx < 2; ++dx)
#pragma unroll
                for (int dy = 0; dy < 2; ++dy) {
                    const int te = (oy + dy) + KT_EY * (ox + dx);
                    const int q = (1 - dy) + 2 * (1 - dx);
#pragma unroll
                    for (int d = 0; d < 3; ++d) lo[d] += sF[3 * q + d][te];
                }
            if (ez >= z0) {
                T *o = out + 3 * NYX * ez + offO;
                if (J0 && (mk & NODE_ACTIVE)) {
                    T *xo = xout + 3 * NYX * ez + offO;
#pragma unroll
                    for (int d = 0; d < 3; ++d) {
                        const T xv = omega * (oR[d] / oD[d]);
                        xo[d] = xv;
                        o[d] = oR[d] - ((mk & (1 << d)) ? xv : lo[d]);
                    }
                } else if (mk & NODE_ACTIVE) {
#pragma unroll
                    for (int d = 0; d < 3; ++d) {
                        T v = (mk & (1 << d)) ? oU[d] : lo[d];
                        if (EPI == EPI_RESID) v = oR[d] - v;
                        if (EPI == EPI_JACOBI) v = oU[d] + (omega * (oR[d] - v)) / oD[d];
                        o[d] = v;
                        if (R64) dsum += (double)v * oW[d];
                        else if (DOT) dsum += (double)v * (double)(EPI == EPI_KU ? oU[d] : oR[d]);
                    }
                } else {
                    o[0] = 0;
                    o[1] = 0;
                    o[2] = 0;
                    if (J0) {
                        T *xo = xout + 3 * NYX * ez + offO;
                        xo[0] = 0;
                        xo[1] = 0;
                        xo[2] = 0;
                    }
                }
            }
        }
        fetchO(ez + 1);
        cur ^= 1;
    }
    if (DOT) {
        // fixed-order workgroup sum (sF is free: every lane has passed its last read of it)
        lds_barrier();
        double *red = reinterpret_cast<double *>(&sF[0][0]);
        red[t] = dsum;
        lds_barrier();
        for (int w = 128; w > 0; w >>= 1) {
            if (t < w) red[t] += red[t + w];
            lds_barrier();
        }
        if (t == 0) partial[tile] = red[0];
    }
}

}  // namespace tg
