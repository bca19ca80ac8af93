// Launch interface of the learned narrow-band extraction (host side).
//
// The lattice's cells are grouped in bricks of 4x4x4 or 8x8x8 cells.  The network predicts
// the (truncated) SDF at every brick centre; a brick whose prediction is
// within its half-diagonal plus a calibrated margin of zero is in the band.
// Band bricks get exact program samples; every other sample takes the
// predicted sign of its brick.  The marching-cubes count and emit kernels
// then run unchanged over the whole lattice (emit reads the stored exact
// samples, so vertex positions are exact).  If no surface cell lies
// in a culled brick, the mesh is bit-identical to the dense extraction.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "mc_kernels.h"
#include "mlp_kernels.h"
#include "scan.h"
#include "sdf_program.h"

namespace iso {

// Brick flags
enum : uint8_t {
  kBrickBand = 1,    // exact samples
  kBrickInside = 2,  // predicted sign (culled bricks): !(f > 0)
  kBrickProven = 4,  // super-brick flags: the centre sample proves the sign
};

// Super-bricks: kCoarse^3 bricks, certified by one centre sample before
// their bricks are.
constexpr int kCoarse = 4;

struct BrickGrid {
  int32_t shift;          // bricks of 2^shift cells per edge
  int32_t brick;          // 2^shift
  int64_t nbx, nby, nbz;  // bricks per axis: ceil(cells / brick)
  int64_t n;              // nbx * nby * nbz
  int64_t ncx, ncy, ncz;  // super-bricks per axis: ceil(bricks / kCoarse)
  int64_t nc;             // ncx * ncy * ncz
};

// shift: kBandBrickShiftDefault or kBandBrickShiftMax.
BrickGrid brick_grid(const Lattice& L, int shift);

// Brick centres (3 floats each, brick order x fastest) and half-diagonals.
hipError_t launch_brick_centres(const Lattice& L, const BrickGrid& G, float* centres,
                                float* half_diag, hipStream_t s);
// Allowances of the Lipschitz certificate beyond its computed bounds: the
// half-diagonal's own f32 rounding (relative) and a fixed absolute margin.
constexpr float kCertifyRelSlack = 1.0e-4f;
constexpr float kCertifyAbsSlack = 2.0e-5f;
// Bound on how far the f32 rounding of two lattice positions (a centre and a
// sample) can move their Euclidean distance, from the lattice's largest
// coordinate magnitude.
float certify_position_slack(const Lattice& L);

// Super-brick certificate: coarse[s] = kBrickProven (| kBrickInside) when
// the program's value at super-brick s's centre is farther from zero than
// the super-brick's half-diagonal plus the bounds on the positions' and the
// evaluations' f32 rounding (sdf_eval_bound), else 0.  G.nc evaluations are
// added to *certified.
hipError_t launch_coarse_certify(const DevOp* ops, int n_ops, const Lattice& L,
                                 const BrickGrid& G, uint8_t* coarse,
                                 unsigned long long* certified, hipStream_t s);

// Brick proofs: proof[b] = its super-brick's flag when that proves the
// sign, else the same test at brick b's own centre (against half_diag[b];
// evaluations added to *certified).
hipError_t launch_brick_certify(const DevOp* ops, int n_ops, const Lattice& L,
                                const BrickGrid& G,
                                const uint8_t* coarse, const float* centres,
                                const float* half_diag, uint8_t* proof,
                                unsigned long long* certified, hipStream_t s);

// Band / inside flags.  With predictions alone, a brick is in the band if
// |pred| <= half_diag + margin, else it takes the predicted sign.  With
// `proof` (launch_brick_certify's flags), a brick leaves the band only when
// its proof allows it, whatever the network predicted, and a brick the
// network culled that its proof does not allow is counted in *overturned.
// The list of band bricks (ascending) and its length go to st.ctl[1].
hipError_t launch_brick_flags(const BrickGrid& G, const float* pred, const float* half_diag,
                              const uint8_t* proof, float margin, uint8_t* flags,
                              uint32_t* band_list, unsigned long long* overturned,
                              const scan::State& st, hipStream_t s);
// Sign bits of the whole lattice from the bricks' predicted signs (and the
// extraction's state reset, like the other signs launchers).
hipError_t launch_band_background(const Lattice& L, const McDims& D, const BrickGrid& G,
                                  const uint8_t* flags, const McBuffers& B, hipStream_t s);
// Exact sign bits of every sample of the band bricks' closed sample boxes,
// each sample evaluated once (by the band brick latest in brick order among
// those holding it); values[f] receives the sample when values is not null.
// n_band bricks from band_list; *evaluated receives the sample count.
hipError_t launch_band_exact(const DevOp* ops, int n_ops, const Lattice& L, const BrickGrid& G,
                             const uint8_t* flags, const uint32_t* band_list, uint64_t n_band,
                             const McBuffers& B, float* values, unsigned long long* evaluated,
                             hipStream_t s);
// Cells of lattice L whose marching-cubes case differs between the sign
// arrays signs_a and signs_b (each with the extraction's guard words); the
// count is added to *differing.
hipError_t launch_band_diff(const Lattice& L, const McDims& D, const uint64_t* signs_a,
                            const uint64_t* signs_b, unsigned long long* differing,
                            hipStream_t s);

}  // namespace iso
