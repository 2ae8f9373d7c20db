// Learned block preconditioner for the local GMRES of the AMEn solver: a
// small MLP that maps, per diagonal block of the local matrix (the blocks
// of the block-Jacobi 'c' preconditioner, tt_linalg.f90:397-429), the
// block's three central diagonals and a residual slice to the block's
// correction.  Device training (forward, backward, Adam) and inference
// kernels, the feature extraction, and the sample collection that records
// (features, block-Jacobi solution) pairs from plain AMEn runs.
#pragma once

#include "amen.h"

#include <cstdint>

namespace ttf {

constexpr int kMlpBlock = 64;               // mode size the network is built for
constexpr int kMlpIn = 4 * kMlpBlock;       // [diag, super, sub, residual]
constexpr int kMlpHid = 256;
constexpr int kMlpOut = kMlpBlock;
// flat parameter layout: W1 (In x Hid) b1 W2 (Hid x Hid) b2 W3 (Hid x Out) b3,
// matrices column-major
constexpr int64_t kMlpOffW1 = 0;
constexpr int64_t kMlpOffB1 = kMlpOffW1 + (int64_t)kMlpIn * kMlpHid;
constexpr int64_t kMlpOffW2 = kMlpOffB1 + kMlpHid;
constexpr int64_t kMlpOffB2 = kMlpOffW2 + (int64_t)kMlpHid * kMlpHid;
constexpr int64_t kMlpOffW3 = kMlpOffB2 + kMlpHid;
constexpr int64_t kMlpOffB3 = kMlpOffW3 + (int64_t)kMlpHid * kMlpOut;
constexpr int64_t kMlpParams = kMlpOffB3 + kMlpOut;

// True once parameters were set (ttf_mlp_init / ttf_mlp_set_params).
bool mlp_ready();

// z = M(v) for the local vector v (rx1, n, rx2) of the operator
// Phi1 A Phi2: features of every block, the network, and the scaled block
// corrections written into z.  n must be kMlpBlock.
void mlp_precond(int rx1, int n, int rx2, int ra1, int ra2, const double* phi1, const double* A,
                 const double* phi2, const double* v, double* z, hipStream_t st);

// When collection is active, records sampled blocks of the pair
// (v, z = block-Jacobi 'c' solution of v) as (features, normalised target).
bool mlp_collecting();
void mlp_collect_pair(int rx1, int n, int rx2, int ra1, int ra2, const double* phi1, const double* A,
                      const double* phi2, const double* v, const double* z, hipStream_t st);

}  // namespace ttf
