// Grid neighbourhood direction tables shared by the generator, the host driver
// and the kernels.
//
// Stored edges run from a voxel towards the negative half of its neighbourhood,
// in the order of the reference's neighbourhood tables:
//   6-connected:  VonNeumannNeighborhood<3>::offset(1..3)  (neighborhood.hpp:51-91)
//   26-connected: MooreNeighborhood<3>::offset(1..13)      (neighborhood.hpp:207-245)
// Directed residual direction d < K/2 is half-offset d; direction d + K/2 is its
// mirror (the positive half). So opposite(d) = (d + K/2) mod K.
#pragma once
#include <stdint.h>

static const int8_t CGP_HALF_OFFSETS_6[3][3] = {
	{-1, 0, 0}, {0, -1, 0}, {0, 0, -1},
};

static const int8_t CGP_HALF_OFFSETS_26[13][3] = {
	{-1, -1, -1}, {0, -1, -1}, {1, -1, -1},
	{-1, 0, -1}, {0, 0, -1}, {1, 0, -1},
	{-1, 1, -1}, {0, 1, -1}, {1, 1, -1},
	{-1, -1, 0}, {0, -1, 0}, {1, -1, 0},
	{-1, 0, 0},
};
