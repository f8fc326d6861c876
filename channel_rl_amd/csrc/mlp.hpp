// mlp.hpp -- parameter layout of the wall-stress network (FP32 master copy):
//   W1 [4][64], b1 [64], W2 [64][64], b2 [64], W3 [64][64], b3 [64],
//   W4 [64][2], b4 [2]      (W[k][n]: input k, output n; 8770 values)
// and the 12 standardisation constants: feature means (4), feature scales
// (4), output means (2), output scales (2).
#pragma once

namespace chan {

enum : int {
    MLP_W1 = 0,
    MLP_B1 = MLP_W1 + 4 * 64,
    MLP_W2 = MLP_B1 + 64,
    MLP_B2 = MLP_W2 + 64 * 64,
    MLP_W3 = MLP_B2 + 64,
    MLP_B3 = MLP_W3 + 64 * 64,
    MLP_W4 = MLP_B3 + 64,
    MLP_B4 = MLP_W4 + 64 * 2,
    MLP_NPARAM = MLP_B4 + 2,
    MLP_NNORM = 12,
};
static_assert(MLP_NPARAM == 8770, "4-64-64-64-2 network");

}  // namespace chan
