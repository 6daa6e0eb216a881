// Tile kernels for one-wave 6-connected tiles (TX * TY == 64 lanes, lane = ly * TX + lx, each lane
// the z-column of TZ voxels) with the tile in registers: the discharge and the global-relabel pass.
//
// Discharge with the whole tile state in registers: excess, residual sink link,
// the 6 residuals and the label of every voxel live in the lane that owns it, x neighbours are
// read across lanes with DPP shifts (row shifts when a tile row is a 16-lane DPP row), y
// neighbours with a lane permute by TX, z neighbours are
// the lane's own registers, and the halo labels (fixed while the tile runs: tiles of one colour
// never touch) are loaded once into the lanes on the tile faces. No LDS, no barriers: a step of
// the push sweep is a few dozen vector instructions instead of a chain of LDS round trips.
//
// The algorithm is the one of k_grid_discharge (the reference's relabel/push loop restricted to a
// tile, include/cugip/advanced_operations/detail/graph_cut_push.hpp:96-186 for the push,
// graph_cut_relabeling.hpp for the distance labels): repeat {region relabel = exact distances to
// t inside the tile given the halo labels; top-down push over the levels that hold excess, each
// voxel of level L pushing first into t, then along its admissible directions in order to level
// L - 1} up to max_iters times, then write back what changed and send the flow that left the tile
// to the neighbouring voxels. What changed is tracked per voxel and residual, generously: a
// sender's admissible residuals and a receiver's four x and y residuals are all written (a store
// of an unchanged value is cheaper than the vector instructions that would tell them apart).
// Included at the end of grid_kernels.h.
#pragma once

namespace cgp {

// Value of lane - 1 (the -x neighbour) or, on the tile's -x face, `edge`.
template <int kTX>
__device__ inline int32_t lane_xm(int32_t v, int32_t edge, int lx)
{
	if constexpr (kTX == 16) {
		// a tile row is a DPP row: lanes with lx == 0 have no source lane and keep `edge`
		return __builtin_amdgcn_update_dpp(edge, v, 0x111 /* row_shr:1 */, 0xf, 0xf, false);
	}
	const int32_t s = __builtin_amdgcn_update_dpp(edge, v, 0x138 /* wave_shr:1 */, 0xf, 0xf, false);
	return lx == 0 ? edge : s;
}

// Value of lane + 1 (the +x neighbour) or, on the +x face, `edge`.
template <int kTX>
__device__ inline int32_t lane_xp(int32_t v, int32_t edge, int lx)
{
	if constexpr (kTX == 16) {
		return __builtin_amdgcn_update_dpp(edge, v, 0x101 /* row_shl:1 */, 0xf, 0xf, false);
	}
	const int32_t s = __builtin_amdgcn_update_dpp(edge, v, 0x130 /* wave_shl:1 */, 0xf, 0xf, false);
	return lx == kTX - 1 ? edge : s;
}

// Value of lane - TX (the -y neighbour) or, on the -y face, `edge`.
template <int kTX>
__device__ inline int32_t lane_ym(int32_t v, int32_t edge, int ly)
{
	const int32_t s = __shfl_up(v, (unsigned)kTX, 64);
	return ly == 0 ? edge : s;
}

// Value of lane + TX (the +y neighbour) or, on the +y face, `edge`.
template <int kTX, int kTY>
__device__ inline int32_t lane_yp(int32_t v, int32_t edge, int ly)
{
	const int32_t s = __shfl_down(v, (unsigned)kTX, 64);
	return ly == kTY - 1 ? edge : s;
}

// Value of lane - 1 (the -x neighbour), or the lane's own value on the tile's -x face.
template <int kTX>
__device__ inline int32_t lane_xm_or_own(int32_t v, int lx)
{
	if constexpr (kTX == 16) {
		return __builtin_amdgcn_update_dpp(v, v, 0x111 /* row_shr:1 */, 0xf, 0xf, false);
	}
	const int32_t s = __builtin_amdgcn_update_dpp(v, v, 0x138 /* wave_shr:1 */, 0xf, 0xf, false);
	return lx == 0 ? v : s;
}

// Value of lane + 1 (the +x neighbour), or the lane's own value on the +x face.
template <int kTX>
__device__ inline int32_t lane_xp_or_own(int32_t v, int lx)
{
	if constexpr (kTX == 16) {
		return __builtin_amdgcn_update_dpp(v, v, 0x101 /* row_shl:1 */, 0xf, 0xf, false);
	}
	const int32_t s = __builtin_amdgcn_update_dpp(v, v, 0x130 /* wave_shl:1 */, 0xf, 0xf, false);
	return lx == kTX - 1 ? v : s;
}

// Region fixpoint h = min(h, 1 + h(nbr)) over the open directions of the lane's column, x and y
// neighbours across lanes (Jacobi rounds), the column swept up and down in registers; halo labels
// fixed, so their terms are applied once before the rounds, and a round reads in-tile neighbours
// only (a lane with no neighbour in a direction reads its own label, which changes nothing).
// Voxels outside the volume have no open direction and keep their label. Labels only fall, so a
// round changed something exactly when the lane's label sum changed. Shared by the relabel pass
// and the discharge's region relabel.
template <int kTX, int kTY, int kTZ>
__device__ inline void wave_relax(int32_t (&lab)[kTZ], const uint32_t (&open)[kTZ], const int32_t (&hx)[kTZ],
                                  const int32_t (&hy)[kTZ], int32_t hzlo, int32_t hzhi, int lx, int ly)
{
	const uint32_t xbit = lx == 0 ? 1u : (lx == kTX - 1 ? 8u : 0u);  // the x face of this lane
	const uint32_t ybit = ly == 0 ? 2u : (ly == kTY - 1 ? 16u : 0u);
	uint32_t sum = 0;
#pragma unroll
	for (int k = 0; k < kTZ; ++k) {
		int32_t b = lab[k];
		b = (open[k] & xbit) && hx[k] + 1 < b ? hx[k] + 1 : b;
		b = (open[k] & ybit) && hy[k] + 1 < b ? hy[k] + 1 : b;
		if (k == 0) {
			b = (open[k] & 4u) && hzlo + 1 < b ? hzlo + 1 : b;
		}
		if (k == kTZ - 1) {
			b = (open[k] & 32u) && hzhi + 1 < b ? hzhi + 1 : b;
		}
		lab[k] = b;
		sum += (uint32_t)b;
	}
	for (;;) {
#pragma unroll
		for (int k = 0; k < kTZ; ++k) {
			const int32_t lp = lab[k] + 1;
			const int32_t xm = lane_xm_or_own<kTX>(lp, lx), xp = lane_xp_or_own<kTX>(lp, lx);
			const int32_t ym = __shfl_up(lp, (unsigned)kTX, 64), yp = __shfl_down(lp, (unsigned)kTX, 64);
			int32_t best = lab[k];
			best = (open[k] & 1u) && xm < best ? xm : best;
			best = (open[k] & 8u) && xp < best ? xp : best;
			best = (open[k] & 2u) && ym < best ? ym : best;
			best = (open[k] & 16u) && yp < best ? yp : best;
			lab[k] = best;
		}
#pragma unroll
		for (int k = 1; k < kTZ; ++k) {
			const int32_t below = lab[k - 1] + 1;
			lab[k] = (open[k] & 4u) && below < lab[k] ? below : lab[k];
		}
#pragma unroll
		for (int k = kTZ - 2; k >= 0; --k) {
			const int32_t above = lab[k + 1] + 1;
			lab[k] = (open[k] & 32u) && above < lab[k] ? above : lab[k];
		}
		uint32_t now = 0;
#pragma unroll
		for (int k = 0; k < kTZ; ++k) {
			now += (uint32_t)lab[k];
		}
		const bool changed = now != sum;
		sum = now;
		if (!__any(changed)) {
			return;
		}
	}
}

// Tiles per workgroup of the one-wave kernels: each wave of a workgroup runs its own tile (no LDS, no
// barriers), the workgroup only packages them. Launches cover every tile of a colour (the list
// lengths stay on the device), and fewer, larger workgroups dispatch faster than many single waves;
// the discharge, whose visits differ most in length, keeps its workgroups small.
constexpr int kDischargeTilesPerBlock = 2, kPassTilesPerBlock = 4;

// Index of this wave's tile in the launch (wave-uniform, so tile addressing stays scalar).
template <int kPerBlock>
__device__ inline int wave_tile_index()
{
	return (int)blockIdx.x * kPerBlock + __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
}

template <int kTX, int kTY, int kTZ>
__global__ __launch_bounds__(64 * kDischargeTilesPerBlock, 4) void k_grid_discharge_wave(GridView gv, const int32_t *__restrict__ list,
                                                            const int32_t *__restrict__ count, int class_stride,
                                                            int max_iters, uint8_t *__restrict__ tile_active)
{
	static_assert(kTX * kTY == 64, "one wave per tile");
	static_assert(kTZ <= 4, "the dirty bits of a column fit one 32-bit word");
	using D = Dirs<6>;
	constexpr int TZ = kTZ;
	// direction bits: 0 -x, 1 -y, 2 -z, 3 +x, 4 +y, 5 +z
	const int tile = discharge_tile(list, count, class_stride, wave_tile_index<kDischargeTilesPerBlock>());
	if (tile < 0) {
		return;
	}
	int local;
	const GridView g = instance_view(gv, tile, local);
	int tx, ty, tz;
	tile_coords(g, local, tx, ty, tz);
	const int ox = tx * kTX, oy = ty * kTY, oz = tz * kTZ;
	const int lane = threadIdx.x & 63;
	const int lx = lane % kTX, ly = lane / kTX;
	const int gx = ox + lx, gy = oy + ly;
	const bool in_xy = gx < g.nx && gy < g.ny;
	const uint32_t xy = slice_xy(g, gx, gy), sxy = (uint32_t)(g.nx * g.ny);

	// ---- loads: own voxels, then the halo labels of the faces this lane lies on
	uint32_t valid = 0;
	int32_t ex[TZ], rt[TZ], r[6][TZ], hx[TZ], hy[TZ];
	const int hxx = lx == 0 ? gx - 1 : gx + 1;  // the x face of this lane (only lanes on one use it)
	const int hyy = ly == 0 ? gy - 1 : gy + 1;
	const bool x_face = (lx == 0 || lx == kTX - 1) && hxx >= 0 && hxx < g.nx && gy < g.ny;
	const bool y_face = (ly == 0 || ly == kTY - 1) && hyy >= 0 && hyy < g.ny && gx < g.nx;
	const uint32_t hxxy = x_face ? slice_xy(g, hxx, gy) : 0u, hyxy = y_face ? slice_xy(g, gx, hyy) : 0u;
#pragma unroll
	for (int k = 0; k < TZ; ++k) {
		const bool v = in_xy && oz + k < g.nz;
		valid |= v ? (1u << k) : 0u;
		ex[k] = v ? tile_at(g.e, g, oz, k, xy, sxy) : 0;
		rt[k] = v ? tile_at(g.rt, g, oz, k, xy, sxy) : 0;
#pragma unroll
		for (int d = 0; d < 6; ++d) {
			r[d][k] = v ? tile_at(g.r + (int64_t)d * g.nvox, g, oz, k, xy, sxy) : 0;
		}
		hx[k] = (x_face && oz + k < g.nz) ? tile_at(g.h, g, oz, k, hxxy, sxy) : kLabelInf;
		hy[k] = (y_face && oz + k < g.nz) ? tile_at(g.h, g, oz, k, hyxy, sxy) : kLabelInf;
	}
	const int32_t hzlo = (in_xy && oz > 0) ? tile_at(g.h, g, oz, -1, xy, sxy) : kLabelInf;
	const int32_t hzhi = (in_xy && oz + TZ < g.nz) ? tile_at(g.h, g, oz, TZ, xy, sxy) : kLabelInf;

	// 6C: flow pushed across the tile faces during the visit (a lane touches at most one x face and
	// one y face, and the two z faces at its column's ends) is the drop of the residual across that
	// face: no flow enters the tile from outside while it runs (the neighbouring tiles are of the
	// other colour), so these residuals only fall. The residuals at load time are kept here.
	const int dfx = lx == 0 ? 0 : 3, dfy = ly == 0 ? 1 : 4;
	const bool on_x = lx == 0 || lx == kTX - 1, on_y = ly == 0 || ly == kTY - 1;
	int32_t out_x[TZ], out_y[TZ];
	const int32_t out_zlo = r[2][0], out_zhi = r[5][TZ - 1];
	uint32_t dirty = 0;  // bits 8k + d: residual d of voxel k changed; bit 8k + 6: excess or rt changed
#pragma unroll
	for (int k = 0; k < TZ; ++k) {
		out_x[k] = on_x ? (dfx == 0 ? r[0][k] : r[3][k]) : 0;
		out_y[k] = on_y ? (dfy == 1 ? r[1][k] : r[4][k]) : 0;
	}
	long long sink = 0;
	int32_t lab[TZ];  // the loaded labels, stored back as they are if no iteration runs
#pragma unroll
	for (int k = 0; k < TZ; ++k) {
		lab[k] = ((valid >> k) & 1u) ? tile_at(g.h, g, oz, k, xy, sxy) : kLabelInf;
	}

	for (int it = 0; it < max_iters; ++it) {
		// ---- region relabel: exact distances to t inside the tile, halo labels fixed
		uint32_t open[TZ];
#pragma unroll
		for (int k = 0; k < TZ; ++k) {
			uint32_t o = 0;
#pragma unroll
			for (int d = 0; d < 6; ++d) {
				o |= r[d][k] > 0 ? (1u << d) : 0u;
			}
			open[k] = o;
			lab[k] = ((valid >> k) & 1u) && rt[k] > 0 ? 1 : kLabelInf;
		}
		wave_relax<kTX, kTY, kTZ>(lab, open, hx, hy, hzlo, hzhi, lx, ly);

		// ---- admissible directions (open, neighbour one level lower), the voxels holding excess
		// and the top level holding excess
		uint32_t adm[TZ];
		uint32_t held = 0;
		int top = 0;
#pragma unroll
		for (int k = 0; k < TZ; ++k) {
			const int32_t l = lab[k];
			const int32_t xm = lane_xm<kTX>(l, hx[k], lx), xp = lane_xp<kTX>(l, hx[k], lx);
			const int32_t ym = lane_ym<kTX>(l, hy[k], ly), yp = lane_yp<kTX, kTY>(l, hy[k], ly);
			const int32_t zm = k == 0 ? hzlo : lab[k > 0 ? k - 1 : 0];
			const int32_t zp = k == TZ - 1 ? hzhi : lab[k < TZ - 1 ? k + 1 : 0];
			uint32_t a = 0;
			a |= xm == l - 1 ? 1u : 0u;
			a |= ym == l - 1 ? 2u : 0u;
			a |= zm == l - 1 ? 4u : 0u;
			a |= xp == l - 1 ? 8u : 0u;
			a |= yp == l - 1 ? 16u : 0u;
			a |= zp == l - 1 ? 32u : 0u;
			// the dirty bits a push sweep step of this voxel sets when it sends: its admissible
			// residuals and its excess (bits 8k + d, 8k + 6), and for the -z / +z directions the
			// receiving voxel's residual back and excess; bits 8k .. 8k + 5 are the admissible set
			a &= open[k];
			uint32_t m = (a | 0x40u) << (8 * k);
			if (k > 0) {
				m |= (a & 4u) ? (0x60u << (8 * (k > 0 ? k - 1 : 0))) : 0u;
			}
			if (k < TZ - 1) {
				m |= (a & 32u) ? (0x44u << (8 * (k < TZ - 1 ? k + 1 : 0))) : 0u;
			}
			adm[k] = m;
			const bool h = ex[k] > 0 && l < kLabelInf;  // (voxels outside the volume hold nothing)
			held |= h ? (1u << k) : 0u;
			top = h && l > top ? l : top;
		}
		int level = wave_max_nonneg(top);
		if (level == 0) {
			break;
		}
		// ---- top-down push sweep over the levels that hold excess: a voxel of level L pushes only
		// to voxels of level L - 1 (exact labels), so the senders of a step never receive in it;
		// the next level is L - 1 if this step pushed inside the tile, else the highest lower level
		// whose voxels held excess when the sweep began. Only level 1 has open sink links (a label
		// is 1 exactly where rt > 0, and rt only falls during the visit) and no admissible
		// direction (no label is 0), so levels >= 2 push along directions and level 1 into t.
		// The levels that held excess, as a wave-uniform mask: bit j for level top - j (j < 64);
		// `deep` when some voxel held excess further below top.
		const int top_level = level;
		uint32_t held_lo = 0, held_hi = 0;
		bool deep_lane = false;
#pragma unroll
		for (int k = 0; k < TZ; ++k) {
			const uint32_t j = (uint32_t)(top_level - lab[k]);
			const bool h = (held >> k) & 1u;
			held_lo |= (h && j < 32u) ? (1u << (j & 31u)) : 0u;
			held_hi |= (h && j >= 32u && j < 64u) ? (1u << (j & 31u)) : 0u;
			deep_lane |= h && j >= 64u;
		}
		const uint64_t held_levels =
		    ((uint64_t)wave_or_dpp(held_hi) << 32) | (uint64_t)wave_or_dpp(held_lo);
		const bool deep = __any(deep_lane);
		while (level >= 2) {
			bool sent = false;
#pragma unroll
			for (int k = 0; k < TZ; ++k) {
				const bool send = lab[k] == level && ex[k] > 0;
				if (!__any(send)) {
					continue;  // no voxel of this z-level sends: nothing to deliver either
				}
				const int32_t e0 = send ? ex[k] : 0;  // non-senders push nothing
				int32_t e = e0;
				int32_t f[6];
#pragma unroll
				for (int d = 0; d < 6; ++d) {
					const int32_t fd = ((adm[k] >> (8 * k + d)) & 1u) ? (e < r[d][k] ? e : r[d][k]) : 0;
					r[d][k] -= fd;
					e -= fd;
					f[d] = fd;
				}
				// a sender's excess and every admissible residual are written back (the directions
				// after the one that used up its excess unchanged), with its z receivers' bits
				dirty |= send ? adm[k] : 0u;
				sent = true;
				// deliveries: the receiver's excess and its residual back towards the sender
				const int32_t from_xm = lane_xm<kTX>(f[3], 0, lx);  // sent +x by lane - 1
				const int32_t from_xp = lane_xp<kTX>(f[0], 0, lx);  // sent -x by lane + 1
				const int32_t from_ym = lane_ym<kTX>(f[4], 0, ly);
				const int32_t from_yp = lane_yp<kTX, kTY>(f[1], 0, ly);
				const int32_t got = from_xm + from_xp + from_ym + from_yp;
				ex[k] += got - (e0 - e);
				r[0][k] += from_xm;
				r[3][k] += from_xp;
				r[1][k] += from_ym;
				r[4][k] += from_yp;
				// a receiver's excess and its four x and y residuals
				dirty |= got > 0 ? (0x5bu << (8 * k)) : 0u;
				// z: the column's own next voxels (pushes across the z faces are counted at the end)
				if (k > 0) {
					ex[k > 0 ? k - 1 : 0] += f[2];
					r[5][k > 0 ? k - 1 : 0] += f[2];
				}
				if (k < TZ - 1) {
					ex[k < TZ - 1 ? k + 1 : 0] += f[5];
					r[2][k < TZ - 1 ? k + 1 : 0] += f[5];
				}
			}
			// after a step with senders, level - 1 is next (it may have received; when all the flow
			// left the tile the step there finds only the voxels that held excess already), else
			// the highest level below that held excess
			if (sent) {
				level = level - 1;
				continue;
			}
			const int j0 = top_level - level + 1;
			const uint64_t below = j0 < 64 ? held_levels >> j0 : 0ull;
			if (below) {
				level = level - 1 - __builtin_ctzll(below);
			} else if (deep) {
				int next = 0;
#pragma unroll
				for (int k = 0; k < TZ; ++k) {
					next = ((held >> k) & 1u) && lab[k] < level && lab[k] > next ? lab[k] : next;
				}
				level = wave_max_nonneg(next);
			} else {
				level = 0;
			}
		}
		if (level == 1) {
#pragma unroll
			for (int k = 0; k < TZ; ++k) {
				const bool send = ((valid >> k) & 1u) && lab[k] == 1 && ex[k] > 0;
				const int32_t fs = send ? (ex[k] < rt[k] ? ex[k] : rt[k]) : 0;
				rt[k] -= fs;
				ex[k] -= fs;
				sink += fs;
				dirty |= fs > 0 ? (1u << (8 * k + 6)) : 0u;
			}
		}
		// when the sweep left no excess in the tile, another region relabel would only find that
		// out; the labels stay valid (pushes along admissible edges keep them valid)
		bool left = false;
#pragma unroll
		for (int k = 0; k < TZ; ++k) {
			left |= ((valid >> k) & 1u) && ex[k] > 0;
		}
		if (!__any(left)) {
			break;
		}
	}

	// ---- write back what changed; pushes that left the tile go to the neighbouring voxels
	bool active = false;
	uint32_t marks = 0;  // bit s: neighbour tile slot s (dz, dy, dx in {-1, 0, 1}) received flow
#pragma unroll
	for (int k = 0; k < TZ; ++k) {
		if (!((valid >> k) & 1u)) {
			continue;
		}
		active |= ex[k] > 0 && lab[k] < kLabelInf;
		tile_at(g.h, g, oz, k, xy, sxy) = lab[k];
		const uint32_t dk = (dirty >> (8 * k)) & 0x7fu;
		if (!dk) {
			continue;
		}
		tile_at(g.e, g, oz, k, xy, sxy) = ex[k];
		tile_at(g.rt, g, oz, k, xy, sxy) = rt[k];
		uint32_t st = (ex[k] > 0 ? kStateExcess : 0u) | (rt[k] > 0 ? kStateSink : 0u);
#pragma unroll
		for (int d = 0; d < 6; ++d) {
			st |= r[d][k] > 0 ? (1u << d) : 0u;
		}
		tile_at(g.omask, g, oz, k, xy, sxy) = st;
#pragma unroll
		for (int d = 0; d < 6; ++d) {
			if ((dk >> d) & 1u) {
				tile_at(g.r + (int64_t)d * g.nvox, g, oz, k, xy, sxy) = r[d][k];
			}
		}
	}
	// flow that crossed the faces goes to the neighbouring voxels (atomics without return)
	auto send_out = [&](int k, int d, int32_t f) {
		if (f <= 0) {
			return;
		}
		const int uk = k + D::dz(d);
		const uint32_t uxy = slice_xy(g, gx + D::dx(d), gy + D::dy(d));
		atomicAdd(&tile_at(g.e, g, oz, uk, uxy, sxy), f);
		atomicAdd(&tile_at(g.r + (int64_t)opposite<6>(d) * g.nvox, g, oz, uk, uxy, sxy), f);
		atomicOr(&tile_at(g.omask, g, oz, uk, uxy, sxy), kStateExcess | (1u << opposite<6>(d)));
		const int sx = D::dx(d) + 1, sy = D::dy(d) + 1, sz = (k + D::dz(d) < 0 ? 0 : (k + D::dz(d) >= TZ ? 2 : 1));
		marks |= 1u << ((sz * 3 + sy) * 3 + sx);
	};
#pragma unroll
	for (int k = 0; k < TZ; ++k) {
		send_out(k, dfx, on_x ? out_x[k] - (dfx == 0 ? r[0][k] : r[3][k]) : 0);
		send_out(k, dfy, on_y ? out_y[k] - (dfy == 1 ? r[1][k] : r[4][k]) : 0);
	}
	send_out(0, 2, out_zlo - r[2][0]);
	send_out(TZ - 1, 5, out_zhi - r[5][TZ - 1]);
	// (a lane on both x faces or both y faces would need two accumulators; TX, TY >= 2 here)
	static_assert(kTX >= 2 && kTY >= 2, "a lane lies on at most one x face and one y face");

	// wave reductions: the flow into t, the activity, the neighbour tiles to mark
#pragma unroll
	for (int off = 32; off > 0; off >>= 1) {
		sink += __shfl_xor(sink, off, 64);
		marks |= (uint32_t)__shfl_xor((int)marks, off, 64);
	}
	const bool any_active = __any(active);
	if (lane == 0) {
		if (sink) {
			atomicAdd(flow_slot(g, tile), (unsigned long long)sink);
		}
		tile_active[tile] = any_active ? 1 : 0;
		atomicAdd(counter_slot(g, 0, tile), 1ull);
	}
	if (lane < 27 && ((marks >> lane) & 1u)) {
		mark_neighbour_tile(g, tile_active, tx, ty, tz, lane);
	}
}

// Global-relabel pass over a list of one-wave 6C tiles, the semantics of k_grid_bfs (labels only
// fall; a tile whose boundary label fell below what a neighbouring voxel could use marks that
// neighbour's tile in dirty_next; the tile's active flag and discharge work estimate are
// refreshed; first_mode 1 / 2 as there) with the tile in registers: own labels and open directions
// in the lanes, the face lanes holding the halo labels.
template <int kTX, int kTY, int kTZ>
__global__ __launch_bounds__(64 * kPassTilesPerBlock) void k_grid_bfs_wave(GridView gv, const int32_t *__restrict__ list,
                                                      const int32_t *__restrict__ count, int32_t threshold,
                                                      int first_mode, uint8_t *__restrict__ dirty_next,
                                                      uint8_t *__restrict__ tile_active,
                                                      uint8_t *__restrict__ tile_work)
{
	static_assert(kTX * kTY == 64 && kTX >= 2 && kTY >= 2, "one wave per tile");
	constexpr int TZ = kTZ;
	const int b = wave_tile_index<kPassTilesPerBlock>();
	if (b >= *count) {
		return;
	}
	const int tile = list[b];
	int local;
	const GridView g = instance_view(gv, tile, local);
	int tx, ty, tz;
	tile_coords(g, local, tx, ty, tz);
	const int ox = tx * kTX, oy = ty * kTY, oz = tz * kTZ;
	const int lane = threadIdx.x & 63;
	const int lx = lane % kTX, ly = lane / kTX;
	const int gx = ox + lx, gy = oy + ly;
	const bool in_xy = gx < g.nx && gy < g.ny;
	const uint32_t xy = slice_xy(g, gx, gy), sxy = (uint32_t)(g.nx * g.ny);
	const int hxx = lx == 0 ? gx - 1 : gx + 1;
	const int hyy = ly == 0 ? gy - 1 : gy + 1;
	// halo labels are read unless this is the first colour's first pass (first_mode 1)
	const bool x_face = first_mode != 1 && (lx == 0 || lx == kTX - 1) && hxx >= 0 && hxx < g.nx && gy < g.ny;
	const bool y_face = first_mode != 1 && (ly == 0 || ly == kTY - 1) && hyy >= 0 && hyy < g.ny && gx < g.nx;
	const uint32_t hxxy = x_face ? slice_xy(g, hxx, gy) : 0u, hyxy = y_face ? slice_xy(g, gx, hyy) : 0u;
	uint32_t valid = 0, excess = 0;
	uint32_t open[TZ];
	int32_t lab[TZ], before[TZ], hx[TZ], hy[TZ];
#pragma unroll
	for (int k = 0; k < TZ; ++k) {
		const bool v = in_xy && oz + k < g.nz;
		valid |= v ? (1u << k) : 0u;
		uint32_t o = 0;
		bool sink_open = false, ex = false;
		if (v) {
			if (threshold == 0) {
				const uint32_t st = tile_at(g.omask, g, oz, k, xy, sxy);
				o = st & 63u;
				ex = (st & kStateExcess) != 0;
				sink_open = (st & kStateSink) != 0;
			} else {
#pragma unroll
				for (int d = 0; d < 6; ++d) {
					o |= tile_at(g.r + (int64_t)d * g.nvox, g, oz, k, xy, sxy) > threshold ? (1u << d) : 0u;
				}
				ex = tile_at(g.e, g, oz, k, xy, sxy) > 0;
				sink_open = tile_at(g.rt, g, oz, k, xy, sxy) > threshold;
			}
		}
		open[k] = o;
		excess |= ex ? (1u << k) : 0u;
		before[k] = (v && first_mode == 0) ? tile_at(g.h, g, oz, k, xy, sxy) : kLabelInf;
		lab[k] = sink_open && before[k] > 1 ? 1 : before[k];
		hx[k] = (x_face && oz + k < g.nz) ? tile_at(g.h, g, oz, k, hxxy, sxy) : kLabelInf;
		hy[k] = (y_face && oz + k < g.nz) ? tile_at(g.h, g, oz, k, hyxy, sxy) : kLabelInf;
	}
	const int32_t hzlo = (first_mode != 1 && in_xy && oz > 0) ? tile_at(g.h, g, oz, -1, xy, sxy) : kLabelInf;
	const int32_t hzhi = (first_mode != 1 && in_xy && oz + TZ < g.nz) ? tile_at(g.h, g, oz, TZ, xy, sxy) : kLabelInf;

	wave_relax<kTX, kTY, kTZ>(lab, open, hx, hy, hzlo, hzhi, lx, ly);

	bool active = false;
	int top = 0, bottom = kLabelInf;
	uint32_t marks = 0;  // bit s: neighbour tile slot s whose voxel could improve
#pragma unroll
	for (int k = 0; k < TZ; ++k) {
		if (!((valid >> k) & 1u)) {
			continue;
		}
		const int32_t nl = lab[k];
		const bool h = ((excess >> k) & 1u) && nl < kLabelInf;
		active |= h;
		top = h && nl > top ? nl : top;
		bottom = nl < bottom ? nl : bottom;
		if (first_mode || nl != before[k]) {
			tile_at(g.h, g, oz, k, xy, sxy) = nl;  // first passes store every label
		}
		if (nl != before[k]) {
			// a voxel beyond the tile pulls this label; it must be relaxed again if it could improve
			if (x_face && hx[k] > nl + 1) {
				marks |= 1u << ((1 * 3 + 1) * 3 + (lx == 0 ? 0 : 2));
			}
			if (y_face && hy[k] > nl + 1) {
				marks |= 1u << ((1 * 3 + (ly == 0 ? 0 : 2)) * 3 + 1);
			}
			if (k == 0 && hzlo > nl + 1 && oz > 0) {
				marks |= 1u << ((0 * 3 + 1) * 3 + 1);
			}
			if (k == TZ - 1 && hzhi > nl + 1 && oz + TZ < g.nz) {
				marks |= 1u << ((2 * 3 + 1) * 3 + 1);
			}
		}
	}
#pragma unroll
	for (int off = 32; off > 0; off >>= 1) {
		marks |= (uint32_t)__shfl_xor((int)marks, off, 64);
	}
	top = wave_max_nonneg(top);
	bottom = wave_min_nonneg(bottom);
	const bool any_active = __any(active);
	if (lane < 27 && ((marks >> lane) & 1u)) {
		mark_neighbour_tile(g, dirty_next, tx, ty, tz, lane);
	}
	if (lane == 0) {
		tile_active[tile] = any_active ? 1 : 0;
		// discharge work estimate: the levels between the tile's lowest label and its highest
		// label holding excess (the discharge runs the heaviest tiles first)
		const int levels = any_active ? top - bottom + 1 : 0;
		tile_work[tile] = (uint8_t)(levels < 255 ? levels : 255);
		atomicAdd(counter_slot(g, 1, tile), 1ull);
	}
}

// S_min closure pass over a list of one-wave 6C tiles, the semantics of k_grid_source_closure
// (s(u) = 1 if u holds excess or some v with s(v) = 1 has residual v -> u > 0; bits only rise;
// first_pass seeds from the excess; the closure is pushed across the tile faces into the
// neighbouring voxels, whose tiles are marked; `reached` is set when a voxel with a residual sink
// link joins) with the tile in registers: a lane's s bits and open directions travel together to
// its x and y neighbour lanes, the column is swept up and down in registers.
template <int kTX, int kTY, int kTZ>
__global__ __launch_bounds__(64 * kPassTilesPerBlock) void k_grid_closure_wave(GridView gv, const int32_t *__restrict__ list,
                                                          const int32_t *__restrict__ count, int first_pass,
                                                          uint8_t *__restrict__ dirty_next,
                                                          int32_t *__restrict__ reached, const SolveControl *ctl)
{
	static_assert(kTX * kTY == 64 && kTX >= 2 && kTY >= 2, "one wave per tile");
	constexpr int TZ = kTZ;
	using D = Dirs<6>;
	const int b = wave_tile_index<kPassTilesPerBlock>();
	if (b >= *count || !gate_open(ctl, kGateClosure | kGateUnlessReached)) {
		return;
	}
	const int tile = list[b];
	int local;
	const GridView g = instance_view(gv, tile, local);
	int tx, ty, tz;
	tile_coords(g, local, tx, ty, tz);
	const int ox = tx * kTX, oy = ty * kTY, oz = tz * kTZ;
	const int lane = threadIdx.x & 63;
	const int lx = lane % kTX, ly = lane / kTX;
	const int gx = ox + lx, gy = oy + ly;
	const bool in_xy = gx < g.nx && gy < g.ny;
	const uint32_t xy = slice_xy(g, gx, gy), sxy = (uint32_t)(g.nx * g.ny);
	uint32_t valid = 0, sbits = 0, before = 0, sink = 0;
	uint32_t open[TZ];
#pragma unroll
	for (int k = 0; k < TZ; ++k) {
		const bool v = in_xy && oz + k < g.nz;
		valid |= v ? (1u << k) : 0u;
		const uint32_t st = v ? tile_at(g.omask, g, oz, k, xy, sxy) : 0u;
		const bool b = v && tile_at(g.s, g, oz, k, xy, sxy) != 0;
		open[k] = st & 63u;
		before |= b ? (1u << k) : 0u;
		sink |= (st & kStateSink) ? (1u << k) : 0u;
		sbits |= (b || (first_pass && (st & kStateExcess))) ? (1u << k) : 0u;
	}
	// fixpoint inside the tile: a lane's s bit with its open directions (bit 6 and bits 0-5) go to
	// the x and y neighbour lanes of the same z-level
	for (;;) {
		const uint32_t old = sbits;
#pragma unroll
		for (int k = 0; k < TZ; ++k) {
			const int32_t pk = (int32_t)(open[k] | (((sbits >> k) & 1u) << 6));
			const int32_t xm = lane_xm<kTX>(pk, 0, lx), xp = lane_xp<kTX>(pk, 0, lx);
			const int32_t ym = lane_ym<kTX>(pk, 0, ly), yp = lane_yp<kTX, kTY>(pk, 0, ly);
			const bool in = ((xm & 64) && (xm & 8)) || ((xp & 64) && (xp & 1)) || ((ym & 64) && (ym & 16)) ||
			                ((yp & 64) && (yp & 2));
			sbits |= (in && ((valid >> k) & 1u)) ? (1u << k) : 0u;
		}
#pragma unroll
		for (int k = 1; k < TZ; ++k) {
			const bool in = ((sbits >> (k - 1)) & 1u) && (open[k - 1] & 32u);
			sbits |= (in && ((valid >> k) & 1u)) ? (1u << k) : 0u;
		}
#pragma unroll
		for (int k = TZ - 2; k >= 0; --k) {
			const bool in = ((sbits >> (k + 1)) & 1u) && (open[k + 1] & 4u);
			sbits |= (in && ((valid >> k) & 1u)) ? (1u << k) : 0u;
		}
		if (!__any(sbits != old)) {
			break;
		}
	}
	// write back own voxels; push the closure across the tile faces
	uint32_t marks = 0;
	auto push_out = [&](int k, int d) {
		uint8_t &su = tile_at(g.s, g, oz, k + D::dz(d), slice_xy(g, gx + D::dx(d), gy + D::dy(d)), sxy);
		if (!su) {
			su = 1;
			const int sz = k + D::dz(d) < 0 ? 0 : (k + D::dz(d) >= TZ ? 2 : 1);
			marks |= 1u << ((sz * 3 + D::dy(d) + 1) * 3 + D::dx(d) + 1);
		}
	};
#pragma unroll
	for (int k = 0; k < TZ; ++k) {
		if (!((sbits >> k) & 1u)) {
			continue;
		}
		if (!((before >> k) & 1u)) {
			tile_at(g.s, g, oz, k, xy, sxy) = 1;
		}
		// (an open direction always leads to a voxel inside the volume)
		if (lx == 0 && (open[k] & 1u)) {
			push_out(k, 0);
		}
		if (lx == kTX - 1 && (open[k] & 8u)) {
			push_out(k, 3);
		}
		if (ly == 0 && (open[k] & 2u)) {
			push_out(k, 1);
		}
		if (ly == kTY - 1 && (open[k] & 16u)) {
			push_out(k, 4);
		}
		if (k == 0 && (open[k] & 4u)) {
			push_out(k, 2);
		}
		if (k == TZ - 1 && (open[k] & 32u)) {
			push_out(k, 5);
		}
	}
	const bool hit = (sbits & sink) != 0;
	if (reached && __any(hit) && lane == 0) {
		*reached = 1;  // every writer stores the same value
	}
#pragma unroll
	for (int off = 32; off > 0; off >>= 1) {
		marks |= (uint32_t)__shfl_xor((int)marks, off, 64);
	}
	if (lane < 27 && ((marks >> lane) & 1u)) {
		mark_neighbour_tile(g, dirty_next, tx, ty, tz, lane);
	}
	if (lane == 0) {
		atomicAdd(counter_slot(g, 2, tile), 1ull);
	}
}

}  // namespace cgp
