// Default grid schedules, one row per (connectivity, batch, grid size class): tile shape,
// discharge phases per global relabel and in-tile iterations per discharge. Written by
// `python -m cugip_amd.autotune --write-header` from timings on an MI355X; policy fields a
// caller sets explicitly override the row. Rows are matched in order: the first row with the
// grid's connectivity and batch flag whose max_voxels (voxels per grid) is not exceeded.
#pragma once

#include <cstdint>

namespace cgp {

struct TunedPolicy {
	int32_t conn, batch;
	int64_t max_voxels;
	int32_t tile_x, tile_y, tile_z;
	int32_t phases, iterations;
};

static const TunedPolicy kTunedPolicies[] = {
	// conn batch max_voxels  tile        phases iters
	{6, 0, 2097152LL, 16, 4, 4, 6, 8},  // grid6_small: 1.00 ms
	{6, 0, INT64_MAX, 32, 2, 4, 8, 32},  // grid6_large: 3.15 ms
	{6, 1, INT64_MAX, 32, 2, 4, 8, 8},  // batch6: 2.36 ms
	{26, 0, 4194304LL, 32, 4, 2, 24, 2},  // grid26_small: 6.88 ms
	{26, 0, INT64_MAX, 32, 4, 2, 24, 2},  // grid26: 10.69 ms
	{26, 1, INT64_MAX, 32, 4, 2, 24, 2},  // batch26: 3.01 ms
};

inline TunedPolicy tuned_policy(int conn, bool batch, int64_t voxels_per_grid)
{
	for (const TunedPolicy &p : kTunedPolicies) {
		if (p.conn == conn && p.batch == (batch ? 1 : 0) && voxels_per_grid <= p.max_voxels) {
			return p;
		}
	}
	return TunedPolicy{conn, batch ? 1 : 0, INT64_MAX, 32, conn == 6 ? 8 : 4, 4, conn == 6 ? 4 : 16, 16};
}

}  // namespace cgp
