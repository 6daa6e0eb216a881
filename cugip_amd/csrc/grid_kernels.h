// HIP kernels of the 3D grid push-relabel (gfx950, wave64). Included by
// grid_solver.hip only.
//
// Algorithm (the reference's relabel/push loop, AOD = include/cugip/advanced_operations/detail):
//   * init:     residuals from capacities (initResidualsKernel, AOD/graph_cut_implementation.hpp:28-40),
//               source t-links saturated into excess (pushThroughTLinksFromSourceKernel,
//               AOD/graph_cut_push.hpp:9-24), plus the per-voxel cancellation min(cs, ct) sent
//               straight to t.
//   * relabel:  exact BFS distance to t over edges with residual > threshold, level 1 = voxels with a
//               residual sink t-link (initBFSKernel + bfsPropagation*, AOD/graph_cut_relabeling.hpp:47-112),
//               computed by tile-local fixpoint relaxation in LDS and passes over dirty tiles.
//   * push:     one top-down level sweep, every voxel of level L pushing min(excess, residual) to any
//               neighbour of lower label and, at level 1, to t (Push::compute + pushImplementation,
//               AOD/graph_cut_push.hpp:160-186, 395-482). Here it runs inside a tile staged in LDS,
//               alternated with a region relabel (exact distances inside the tile given the halo
//               labels), so that flow crosses a whole tile per HBM round trip. Neighbouring tiles
//               never run together (colour phases), so the only cross-tile write that needs an
//               atomic is the excess of a halo voxel.
//   * cut:      T_min = finite labels of the last exact relabel (fill_segments, graph_cut.hpp:304-333);
//               S_min = closure of the excess voxels under residual edges (see DESIGN.md proof).
#pragma once

#include "common.h"
#include "grid_control.h"
#include "grid_tile.h"

namespace cgp {

// Per-voxel state word (GridView::omask): bit d set while the residual in direction d is positive,
// plus whether the residual sink link and the excess are positive. The init kernel writes it,
// the discharge keeps it current (its own voxels at write-back, atomic ORs into the halo voxels it
// pushes to), and the relabel and closure kernels read it instead of the K residual planes.
constexpr uint32_t kStateSink = 1u << 29;
constexpr uint32_t kStateExcess = 1u << 30;

// Tile of discharge block b. A contiguous list (class_stride 0) has its length in count[0]; a list
// in class regions (class_stride > 0) keeps class c's tiles at list + c * class_stride and their
// number in count[c], and blocks take the classes in order. -1: no tile for this block.
__device__ inline int discharge_tile(const int32_t *list, const int32_t *count, int class_stride, int b)
{
	if (class_stride == 0) {
		return b < *count ? list[b] : -1;
	}
	int base = 0;
#pragma unroll
	for (int c = 0; c < kWorkClasses; ++c) {
		const int n = count[c];
		if (b < base + n) {
			return list[c * class_stride + (b - base)];
		}
		base += n;
	}
	return -1;
}

// ---------------------------------------------------------------------------------------------
// init: residual planes, cancellation, excess, residual sink link. Sets status bit 0 for a negative
// capacity, bit 1 when a voxel's inflow or an edge's capacity pair does not fit int32, and bit 2
// when some edge's capacity pair exceeds 32767 (then residuals cannot be staged as int16).
// Voxel-parallel over z-columns rather than tiles: the xy plane of a grid is cut into strips of
// kInitThreads consecutive xy indices (x fastest, running on across rows, so no lane idles on a
// partial tile), z into chunks of kInitDepth; workgroup b takes one strip × one chunk (strip-major
// within its chunk; block_inst / inst_first_block give the grid of b in a batch), thread t the
// column of its xy index, whose +z neighbours are its own next voxels. 6C: all loads of a voxel are
// issued before its arithmetic and stores; 26C: one residual plane at a time over the column.
// 26C instead cuts the xy plane into kInitTileX x kInitTileY rectangles, so that the columns whose
// capacities a column's mirror residuals read (its 8 xy neighbours) mostly belong to the same
// workgroup.
constexpr int kInitThreads = 256, kInitDepth = 8, kInitTileX = 64, kInitTileY = kInitThreads / kInitTileX;

__host__ __device__ inline int64_t init_blocks(int64_t nx, int64_t ny, int64_t nz, int conn)
{
	const int64_t per_chunk = conn == 26 ? ((nx + kInitTileX - 1) / kInitTileX) * ((ny + kInitTileY - 1) / kInitTileY)
	                                     : (nx * ny + kInitThreads - 1) / kInitThreads;
	return per_chunk * ((nz + kInitDepth - 1) / kInitDepth);
}

// Column (x, y) and z-chunk of thread threadIdx.x of init workgroup lb in the xy-tile layout.
__device__ inline void init_tile_coords(const GridView &g, int lb, int &x, int &y, int &chunk)
{
	const int tiles_x = (g.nx + kInitTileX - 1) / kInitTileX;
	const int tiles = tiles_x * ((g.ny + kInitTileY - 1) / kInitTileY);
	chunk = lb / tiles;
	const int t = lb - chunk * tiles;
	const int ty = t / tiles_x, tx = t - ty * tiles_x;
	x = tx * kInitTileX + (int)(threadIdx.x % kInitTileX);
	y = ty * kInitTileY + (int)(threadIdx.x / kInitTileX);
}

// The column work of k_grid_init one capacity plane at a time: plane h gives residual plane h of
// every voxel of the column (the pair's capacity stored at the voxel) and its mirror plane h + H
// (the capacity stored at the neighbour v - offset(h)), both loaded together so that the
// neighbour's column, which a nearby thread of the workgroup loads for its own plane h, is read
// from the cache. Only one plane's address is live: voxel by voxel, all 39 plane addresses of a
// 26-connected grid would stay live across the column, more than the scalar registers hold.
// State bits and inflow sums stay in registers until the last plane.
template <int kConn, bool kBack>
__device__ inline void init_column_by_plane(const GridView &g, int x, int y, int oz, uint32_t xy, uint32_t sxy,
                                            long long &cancelled, bool &negative, bool &overflow, bool &wide)
{
	using D = Dirs<kConn>;
	constexpr int H = D::H, Z = kInitDepth;
	// inflow sums saturate at 2^31 (each term is below 2^31, so the 32-bit sum cannot wrap): only
	// whether one exceeds int32 matters; likewise a capacity pair's sum of two values below 2^31
	uint32_t state[Z], incoming[Z];
#pragma unroll
	for (int k = 0; k < Z; ++k) {
		state[k] = 0;
		incoming[k] = 0;
		if (oz + k >= g.nz) {
			continue;
		}
		const int32_t cs = tile_at(g.cs, g, oz, k, xy, sxy), ct = tile_at(g.ct, g, oz, k, xy, sxy);
		negative |= (cs < 0) | (ct < 0);
		const int32_t m = cs < ct ? cs : ct;
		cancelled += m;
		incoming[k] = (uint32_t)(cs - m);
		state[k] = (cs - m > 0 ? kStateExcess : 0u) | (ct - m > 0 ? kStateSink : 0u);
		tile_at(g.e, g, oz, k, xy, sxy) = cs - m;
		tile_at(g.rt, g, oz, k, xy, sxy) = ct - m;
	}
#pragma unroll 1
	for (int h = 0; h < H; ++h) {
		// direction h leads to v + offset(h); its mirror h + H to u = v - offset(h), whose own
		// plane-h capacity pair is the edge (u, v)
		const int dx = D::dx(h), dy = D::dy(h), dz = D::dz(h);
		const bool own_xy = x + dx >= 0 && x + dx < g.nx && y + dy >= 0 && y + dy < g.ny;
		const bool mir_xy = x - dx >= 0 && x - dx < g.nx && y - dy >= 0 && y - dy < g.ny;
		const int32_t *pn = g.cn + (int64_t)h * g.nvox;
		const int32_t *pb = kBack ? g.cnb + (int64_t)h * g.nvox : nullptr;
		// every lane loads (from its own column where u lies outside the grid in x or y) and keeps
		// the value only for an edge that exists: no divergent loads; the z tests are uniform
		const uint32_t uxy = mir_xy ? slice_xy(g, x - dx, y - dy) : xy;
		// the column's offset made opaque to the compiler inside each plane's step, so that its
		// per-slice offsets are not hoisted out of the loop as 64-bit values (two registers each
		// across all planes) but formed next to each load and store as 32-bit offsets
		uint32_t vxy = xy;
		asm volatile("" : "+v"(vxy));
		int32_t out[Z], in[Z], mout[Z], min_[Z];
#pragma unroll
		for (int k = 0; k < Z; ++k) {
			out[k] = in[k] = mout[k] = min_[k] = 0;
			const int z = oz + k;
			if (z >= g.nz) {
				continue;
			}
			const bool own = own_xy && (unsigned)(z + dz) < (unsigned)g.nz;
			const int32_t c = tile_at(pn, g, oz, k, vxy, sxy);
			const int32_t b = kBack ? tile_at(pb, g, oz, k, vxy, sxy) : c;
			out[k] = own ? c : 0;
			in[k] = own ? b : 0;
			// u's slice, or the voxel's own one where u lies below or above the grid
			const bool mz = (unsigned)(z - dz) < (unsigned)g.nz;
			const int64_t ms = slice_start(g, mz ? z - dz : z);
			const int32_t mc = at(pn + ms, uxy);
			const int32_t mb = kBack ? at(pb + ms, uxy) : mc;
			min_[k] = mir_xy && mz ? mc : 0;
			mout[k] = mir_xy && mz ? mb : 0;
		}
		int32_t *pr = g.r + (int64_t)h * g.nvox, *pm = g.r + (int64_t)(h + H) * g.nvox;
#pragma unroll
		for (int k = 0; k < Z; ++k) {
			if (oz + k >= g.nz) {
				continue;
			}
			negative |= (out[k] | in[k] | mout[k] | min_[k]) < 0;
			const uint32_t pair = (uint32_t)out[k] + (uint32_t)in[k];
			const uint32_t mpair = (uint32_t)mout[k] + (uint32_t)min_[k];
			overflow |= (pair | mpair) > 0x7fffffffu;
			wide |= (pair | mpair) > 0x7fffu;
			incoming[k] = min(incoming[k] + (uint32_t)in[k], 0x80000000u);
			incoming[k] = min(incoming[k] + (uint32_t)min_[k], 0x80000000u);
			tile_at(pr, g, oz, k, vxy, sxy) = out[k];
			tile_at(pm, g, oz, k, vxy, sxy) = mout[k];
			state[k] |= (out[k] > 0 ? (1u << h) : 0u) | (mout[k] > 0 ? (1u << (h + H)) : 0u);
		}
	}
#pragma unroll
	for (int k = 0; k < Z; ++k) {
		if (oz + k >= g.nz) {
			continue;
		}
		overflow |= incoming[k] > 0x7fffffffu;
		tile_at(g.omask, g, oz, k, xy, sxy) = state[k];
	}
}

template <int kConn>
__global__ __launch_bounds__(kInitThreads) void k_grid_init(GridView gv, const int32_t *__restrict__ block_inst,
                                                           const int32_t *__restrict__ inst_first_block,
                                                           const SolveControl *ctl)
{
	using D = Dirs<kConn>;
	constexpr int K = D::K, H = D::H;
	__shared__ long long red[kInitThreads / 64];
	if (!gate_open(ctl, kGateInit)) {
		return;
	}
	int lb = blockIdx.x;
	GridView g = gv;
	if (gv.inst) {
		const int i = block_inst[blockIdx.x];
		lb = blockIdx.x - inst_first_block[i];
		g = instance_at(gv, i);
	}
	const uint32_t sxy = (uint32_t)(g.nx * g.ny);
	long long cancelled = 0;
	bool negative = false, overflow = false, wide = false;
	// 6C: the whole column unrolled, voxel by voxel; 26C: plane by plane
	constexpr int kUnroll = kInitDepth;
	if constexpr (kConn == 26) {
		int x, y, chunk;
		init_tile_coords(g, lb, x, y, chunk);
		if (x < g.nx && y < g.ny) {
			if (g.cnb) {
				init_column_by_plane<kConn, true>(g, x, y, chunk * kInitDepth, slice_xy(g, x, y), sxy, cancelled,
				                                  negative, overflow, wide);
			} else {
				init_column_by_plane<kConn, false>(g, x, y, chunk * kInitDepth, slice_xy(g, x, y), sxy, cancelled,
				                                   negative, overflow, wide);
			}
		}
	} else {
		const int strips = (int)((sxy + kInitThreads - 1) / kInitThreads);
		const int chunk = lb / strips, strip = lb - chunk * strips;
		const uint32_t xy = (uint32_t)strip * kInitThreads + threadIdx.x;
		const int y = (int)(xy / (uint32_t)g.nx), x = (int)(xy - (uint32_t)y * (uint32_t)g.nx);
		const int oz = chunk * kInitDepth;
		if (xy < sxy) {
#pragma unroll kUnroll
			for (int k = 0; k < kInitDepth; ++k) {
				const int z = oz + k;
				if (z >= g.nz) {
					continue;
				}
				int32_t out[K], in[K];
				const int32_t cs = tile_at(g.cs, g, oz, k, xy, sxy), ct = tile_at(g.ct, g, oz, k, xy, sxy);
#pragma unroll
				for (int d = 0; d < K; ++d) {
					out[d] = in[d] = 0;
					if (!inside(g, x + D::dx(d), y + D::dy(d), z + D::dz(d))) {
						continue;
					}
					if (d < H) {
						out[d] = tile_at(g.cn + (int64_t)d * g.nvox, g, oz, k, xy, sxy);
						in[d] = g.cnb ? tile_at(g.cnb + (int64_t)d * g.nvox, g, oz, k, xy, sxy) : out[d];
					} else {
						const uint32_t uxy = slice_xy(g, x + D::dx(d), y + D::dy(d));
						const int uk = k + D::dz(d);
						in[d] = tile_at(g.cn + (int64_t)(d - H) * g.nvox, g, oz, uk, uxy, sxy);
						out[d] = g.cnb ? tile_at(g.cnb + (int64_t)(d - H) * g.nvox, g, oz, uk, uxy, sxy) : in[d];
					}
				}
				negative |= (cs < 0) | (ct < 0);
				const int32_t m = cs < ct ? cs : ct;
				cancelled += m;
				long long incoming = cs - m;
				uint32_t state = (cs - m > 0 ? kStateExcess : 0u) | (ct - m > 0 ? kStateSink : 0u);
#pragma unroll
				for (int d = 0; d < K; ++d) {
					negative |= (out[d] < 0) | (in[d] < 0);
					overflow |= ((long long)out[d] + (long long)in[d]) > 0x7fffffffLL;
					wide |= ((long long)out[d] + (long long)in[d]) > 0x7fffLL;
					incoming += in[d];
					tile_at(g.r + (int64_t)d * g.nvox, g, oz, k, xy, sxy) = out[d];
					state |= out[d] > 0 ? (1u << d) : 0u;
				}
				overflow |= incoming > 0x7fffffffLL;
				tile_at(g.omask, g, oz, k, xy, sxy) = state;
				tile_at(g.e, g, oz, k, xy, sxy) = cs - m;
				tile_at(g.rt, g, oz, k, xy, sxy) = ct - m;
			}
		}
	}
	const int bits = (negative ? 1 : 0) | (overflow ? 2 : 0) | (wide ? 4 : 0);
	// (a plain read first: when every workgroup sets a bit, one atomic each on one word would
	// serialise in the L2)
	if (bits && (*(volatile int32_t *)g.status & bits) != bits) {
		atomicOr(g.status, bits);
	}
	const long long total = block_sum_i64<kInitThreads>(cancelled, red);
	if (threadIdx.x == 0 && total) {
		atomicAdd(flow_slot(g, blockIdx.x), (unsigned long long)total);
	}
}

// ---------------------------------------------------------------------------------------------
// Stage the label tile with its halo into LDS; voxels outside the volume read as kLabelInf. Thread
// (lx, ly) loads its own column with the two z-halo voxels (one z-slice per load, so the slice
// start is uniform), and the threads share the lateral halo ring of every z-level. All loads are
// issued before the first LDS store, so that one memory latency covers them all.
template <typename T>
struct HaloLabels {
	static constexpr int kRing = 2 * T::PX + 2 * (T::PY - 2);  // lateral halo voxels of one z-level
	static constexpr int kRingPer = (kRing * T::PZ + T::NT - 1) / T::NT;
	int32_t col[T::PZ];
	int32_t ring[kRingPer];

	// halo-box position of ring entry r of a z-level
	__device__ static void ring_pos(int r, int &px, int &py)
	{
		if (r < T::PX) {
			px = r;
			py = 0;
		} else if (r < 2 * T::PX) {
			px = r - T::PX;
			py = T::PY - 1;
		} else {
			const int q = r - 2 * T::PX;
			px = (q & 1) ? T::PX - 1 : 0;
			py = 1 + q / 2;
		}
	}
};

template <typename T>
__device__ inline void fetch_label_tile(const GridView &g, HaloLabels<T> &halo, int ox, int oy, int oz)
{
	const int tid = threadIdx.x;
	const int gx = ox + tid % T::TX, gy = oy + tid / T::TX;
	const bool in_xy = gx < g.nx && gy < g.ny;
	// every halo voxel lies at or above slice zb: 32-bit offsets from there
	const int zb = oz > 0 ? oz - 1 : 0;
	const int32_t *base = g.h + slice_start(g, zb);
	const uint32_t xy = slice_xy(g, gx, gy), sxy = (uint32_t)(g.nx * g.ny);
#pragma unroll
	for (int pz = 0; pz < T::PZ; ++pz) {
		const int z = oz + pz - 1;
		int32_t v = kLabelInf;
		if (in_xy && z >= 0 && z < g.nz) {
			v = at(base, (uint32_t)(z - zb) * sxy + xy);
		}
		halo.col[pz] = v;
	}
#pragma unroll
	for (int it = 0; it < HaloLabels<T>::kRingPer; ++it) {
		const int e = tid + it * T::NT;
		int32_t v = kLabelInf;
		if (e < HaloLabels<T>::kRing * T::PZ) {
			const int pz = e / HaloLabels<T>::kRing;
			int px, py;
			HaloLabels<T>::ring_pos(e % HaloLabels<T>::kRing, px, py);
			const int x = ox + px - 1, y = oy + py - 1, z = oz + pz - 1;
			if (inside(g, x, y, z)) {
				v = at(base, (uint32_t)(z - zb) * sxy + slice_xy(g, x, y));
			}
		}
		halo.ring[it] = v;
	}
}

template <typename T>
__device__ inline void store_label_tile(int32_t *hl, const HaloLabels<T> &halo)
{
	const int tid = threadIdx.x;
	const int lx = tid % T::TX, ly = tid / T::TX;
#pragma unroll
	for (int pz = 0; pz < T::PZ; ++pz) {
		hl[T::pidx(lx, ly, pz - 1)] = halo.col[pz];
	}
#pragma unroll
	for (int it = 0; it < HaloLabels<T>::kRingPer; ++it) {
		const int e = tid + it * T::NT;
		if (e < HaloLabels<T>::kRing * T::PZ) {
			const int pz = e / HaloLabels<T>::kRing;
			int px, py;
			HaloLabels<T>::ring_pos(e % HaloLabels<T>::kRing, px, py);
			hl[(pz * T::PY + py) * T::PX + px] = halo.ring[it];
		}
	}
}

// Mark the neighbouring tile in direction (dx, dy, dz) of tile (tx, ty, tz), if it exists.
__device__ inline void mark_neighbour_tile(const GridView &g, uint8_t *flags, int tx, int ty, int tz, int slot)
{
	const int dx = slot % 3 - 1, dy = (slot / 3) % 3 - 1, dz = slot / 9 - 1;
	const int nx = tx + dx, ny = ty + dy, nz = tz + dz;
	if (nx >= 0 && nx < g.ntx && ny >= 0 && ny < g.nty && nz >= 0 && nz < g.ntz) {
		flags[g.tbase + (nz * g.nty + ny) * g.ntx + nx] = 1;
	}
}

// Index of the direction with offset (x, y, z), -1 if the neighbourhood has none.
template <typename D>
__host__ __device__ constexpr int dir_index(int x, int y, int z)
{
	for (int d = 0; d < D::K; ++d) {
		if (D::dx(d) == x && D::dy(d) == y && D::dz(d) == z) {
			return d;
		}
	}
	return -1;
}

// Fixpoint of h(u) = min(h(u), 1 + h(v)) over the open directions u -> v of the tile's valid
// voxels, the halo labels fixed; hl holds the result when it returns. A round pulls the
// directions that leave the thread's z-column from LDS and then runs the column up and down in
// registers; the halo ends of the column never change and are read once. The fixpoint does not
// depend on the order of the relaxations.
template <typename T>
__device__ inline void relax_tile(int32_t *hl, uint32_t (&open)[T::TZ], uint32_t valid, int lx, int ly)
{
	using D = typename T::D;
	constexpr int kMZ = dir_index<D>(0, 0, -1), kPZ = dir_index<D>(0, 0, 1);
	int32_t lab[T::TZ], stored[T::TZ];
#pragma unroll
	for (int k = 0; k < T::TZ; ++k) {
		lab[k] = stored[k] = hl[T::pidx(lx, ly, k)];
	}
	const int32_t below = hl[T::pidx(lx, ly, -1)] + 1, above = hl[T::pidx(lx, ly, T::TZ)] + 1;
	for (;;) {
		// open[] is left as it is; the empty asm keeps the optimiser from hoisting per-direction
		// lane masks out of the round loop into scalar registers
#pragma unroll
		for (int k = 0; k < T::TZ; ++k) {
			asm volatile("" : "+v"(open[k]));
		}
#pragma unroll
		for (int k = 0; k < T::TZ; ++k) {
			if (!((valid >> k) & 1u)) {
				continue;
			}
			const int p = T::pidx(lx, ly, k);
			int32_t best = lab[k];
#pragma unroll
			for (int d = 0; d < T::K; ++d) {
				if ((D::dx(d) != 0 || D::dy(d) != 0) && (open[k] & (1u << d))) {
					const int32_t nb = hl[p + T::poff(d)] + 1;
					best = nb < best ? nb : best;
				}
			}
			if (k == 0 && (open[k] & (1u << kMZ)) && below < best) {
				best = below;
			}
			if (k == T::TZ - 1 && (open[k] & (1u << kPZ)) && above < best) {
				best = above;
			}
			lab[k] = best;
		}
#pragma unroll
		for (int k = 1; k < T::TZ; ++k) {
			if (((valid >> k) & 1u) && ((open[k] >> kMZ) & 1u) && lab[k - 1] + 1 < lab[k]) {
				lab[k] = lab[k - 1] + 1;
			}
		}
#pragma unroll
		for (int k = T::TZ - 2; k >= 0; --k) {
			if (((valid >> k) & 1u) && ((open[k] >> kPZ) & 1u) && lab[k + 1] + 1 < lab[k]) {
				lab[k] = lab[k + 1] + 1;
			}
		}
		int changed = 0;
#pragma unroll
		for (int k = 0; k < T::TZ; ++k) {
			if (lab[k] < stored[k]) {
				hl[T::pidx(lx, ly, k)] = lab[k];
				stored[k] = lab[k];
				changed = 1;
			}
		}
		if (!__syncthreads_or(changed)) {
			break;
		}
	}
}

// Neighbour-tile slot (0..26) of the voxel beyond own voxel (lx, ly, lz) in direction d,
// or -1 when that voxel lies inside the tile.
template <typename T>
__device__ inline int leaving_slot(int lx, int ly, int lz, int d)
{
	using D = typename T::D;
	const int x = lx + D::dx(d), y = ly + D::dy(d), z = lz + D::dz(d);
	const int sx = x < 0 ? 0 : (x >= T::TX ? 2 : 1);
	const int sy = y < 0 ? 0 : (y >= T::TY ? 2 : 1);
	const int sz = z < 0 ? 0 : (z >= T::TZ ? 2 : 1);
	const int slot = (sz * 3 + sy) * 3 + sx;
	return slot == 13 ? -1 : slot;
}

// ---------------------------------------------------------------------------------------------
// Global relabel pass over a list of tiles. Labels start at kLabelInf (a fill before the first
// pass) and only fall: level 1 = residual sink link > threshold, then 1 + the lowest neighbour
// reachable over an edge with residual > threshold. Tiles run concurrently and may read stale
// halo labels; a tile whose boundary label fell below what a neighbouring voxel could use marks
// that neighbour's tile in dirty_next, so a pass that marks nothing ends the relabel. Every
// processed tile refreshes its active flag (some voxel with excess and a finite label).
// first_mode replaces the fill on 6-connected grids, whose halo holds only face neighbours, all of
// the other sweep colour: 1 = first pass of the first colour (own and halo labels read as
// kLabelInf, nothing is loaded), 2 = first pass of the second colour (own labels read as kLabelInf,
// the halo is the first colour's fresh labels); both store every own label.
template <int kConn, int kTX, int kTY, int kTZ>
__global__ __launch_bounds__(kTX *kTY) void k_grid_bfs(GridView gv, const int32_t *__restrict__ list,
                                                       const int32_t *__restrict__ count, int32_t threshold,
                                                       int first_mode, uint8_t *__restrict__ dirty_next,
                                                       uint8_t *__restrict__ tile_active,
                                                       uint8_t *__restrict__ tile_work)
{
	using T = Tile<kConn, kTX, kTY, kTZ>;
	const int tile = list[blockIdx.x];  // in bounds: the grid never exceeds the list's capacity
	if ((int)blockIdx.x >= *count) {
		return;
	}
	int local;
	const GridView g = instance_view(gv, tile, local);
	int tx, ty, tz;
	tile_coords(g, local, tx, ty, tz);
	const int ox = tx * T::TX, oy = ty * T::TY, oz = tz * T::TZ;

	__shared__ int32_t hl[T::NP];
	__shared__ uint8_t marks[27];
	__shared__ int span[2];  // highest label holding excess, lowest finite label
	const int tid = threadIdx.x;
	const int lx = tid % T::TX, ly = tid / T::TX;
	const int gx = ox + lx, gy = oy + ly;
	if (tid < 27) {
		marks[tid] = 0;
	}
	if (tid == 0) {
		span[0] = 0;
		span[1] = kLabelInf;
	}
	// every global load of the visit first (halo labels, state words), then the LDS stores
	HaloLabels<T> halo;
	if (first_mode == 1) {
#pragma unroll
		for (int pz = 0; pz < T::PZ; ++pz) {
			halo.col[pz] = kLabelInf;
		}
#pragma unroll
		for (int it = 0; it < HaloLabels<T>::kRingPer; ++it) {
			halo.ring[it] = kLabelInf;
		}
	} else {
		fetch_label_tile<T>(g, halo, ox, oy, oz);
		if (first_mode == 2) {
#pragma unroll
			for (int k = 0; k < T::TZ; ++k) {
				halo.col[k + 1] = kLabelInf;  // own labels are not written yet
			}
		}
	}
	const uint32_t xy = slice_xy(g, gx, gy), sxy = (uint32_t)(g.nx * g.ny);
	uint32_t valid = 0;
	uint32_t state[T::TZ];
#pragma unroll
	for (int k = 0; k < T::TZ; ++k) {
		valid |= (gx < g.nx && gy < g.ny && oz + k < g.nz) ? (1u << k) : 0u;
		state[k] = (threshold == 0 && (valid & (1u << k))) ? tile_at(g.omask, g, oz, k, xy, sxy) : 0u;
	}
	store_label_tile<T>(hl, halo);
	__syncthreads();

	uint32_t open[T::TZ];
	int32_t before[T::TZ];
	bool has_excess[T::TZ];
#pragma unroll
	for (int k = 0; k < T::TZ; ++k) {
		open[k] = 0;
		has_excess[k] = false;
		const int p = T::pidx(lx, ly, k);
		before[k] = hl[p];
		if (valid & (1u << k)) {
			const int64_t v = voxel_index(g, gx, gy, oz + k);
			bool sink_open;
			if (threshold == 0) {
				// exact relabel: the state word has every open direction
				const uint32_t st = state[k];
				open[k] = st & ((1u << T::K) - 1u);
				has_excess[k] = (st & kStateExcess) != 0;
				sink_open = (st & kStateSink) != 0;
			} else {
#pragma unroll
				for (int d = 0; d < T::K; ++d) {
					open[k] |= (g.r[(int64_t)d * g.nvox + v] > threshold) ? (1u << d) : 0u;
				}
				has_excess[k] = g.e[v] > 0;
				sink_open = g.rt[v] > threshold;
			}
			if (sink_open && hl[p] > 1) {
				hl[p] = 1;
			}
		}
	}
	__syncthreads();
	relax_tile<T>(hl, open, valid, lx, ly);

	int active = 0;
	int top = 0, bottom = kLabelInf;
#pragma unroll
	for (int k = 0; k < T::TZ; ++k) {
		if (!(valid & (1u << k))) {
			continue;
		}
		const int p = T::pidx(lx, ly, k);
		const int32_t nl = hl[p];
		active |= has_excess[k] && nl < kLabelInf;
		if (has_excess[k] && nl < kLabelInf && nl > top) {
			top = nl;
		}
		bottom = nl < bottom ? nl : bottom;
		if (first_mode && nl == before[k]) {
			tile_at(g.h, g, oz, k, xy, sxy) = nl;  // kLabelInf over whatever the last solve left
		}
		if (nl != before[k]) {
			tile_at(g.h, g, oz, k, xy, sxy) = nl;
#pragma unroll
			for (int d = 0; d < T::K; ++d) {
				// a voxel beyond the tile pulls this label; it must be relaxed again if it could improve
				const int slot = leaving_slot<T>(lx, ly, k, d);
				if (slot >= 0 && hl[p + T::poff(d)] > nl + 1) {
					marks[slot] = 1;
				}
			}
		}
	}
	top = wave_max_i32(top);
	bottom = -wave_max_i32(-bottom);
	if ((tid & 63) == 0) {
		atomicMax(&span[0], top);
		atomicMin(&span[1], bottom);
	}
	active = __syncthreads_or(active);
	if (tid < 27 && marks[tid]) {
		mark_neighbour_tile(g, dirty_next, tx, ty, tz, tid);
	}
	if (tid == 0) {
		tile_active[tile] = active ? 1 : 0;
		// discharge work estimate: the levels between the tile's lowest label and its highest
		// label holding excess (the discharge runs the heaviest tiles first)
		const int levels = active ? span[0] - span[1] + 1 : 0;
		tile_work[tile] = (uint8_t)(levels < 255 ? levels : 255);
		atomicAdd(counter_slot(g, 1, tile), 1ull);
	}
}

// Admissible direction masks of a thread's TZ voxels: packed into one word when K * TZ <= 32.
template <int K, int TZ>
struct AdmissibleDirs {
	static constexpr bool kPacked = K * TZ <= 32;
	uint32_t w[kPacked ? 1 : TZ] = {};
	__device__ void set(int k, uint32_t m)
	{
		if (kPacked) {
			w[0] = (w[0] & ~(((1u << K) - 1u) << (K * k))) | (m << (K * k));
		} else {
			w[kPacked ? 0 : k] = m;
		}
	}
	__device__ uint32_t get(int k) const
	{
		return kPacked ? (w[0] >> (K * k)) & ((1u << K) - 1u) : w[kPacked ? 0 : k];
	}
};

// ---------------------------------------------------------------------------------------------
// Discharge one tile: repeated {region relabel, top-down level push} in LDS, then write back.
// Runs only on tiles of one colour at a time, so no neighbouring tile is running.
// Residuals are staged in LDS as RT: int16_t when every edge's capacity pair fits (a residual never
// exceeds it), int32_t otherwise. For 6-connected grids the directions written during the visit
// are kept as one byte per voxel, four to a 32-bit word, so that LDS atomics can set them.
template <int kConn>
struct DirtyBits {
	static constexpr int kWords(int nv) { return kConn <= 8 ? (nv + 3) / 4 : nv; }
	__device__ static void set(uint32_t *dm, int i, int d)
	{
		if (kConn <= 8) {
			// the empty asm keeps the optimiser from hoisting the word address and bit of every
			// (voxel, direction) pair out of the push loop, where they would not fit in registers
			asm volatile("" : "+v"(i));
			atomicOr(&dm[i >> 2], 1u << (8 * (i & 3) + d));
		} else {
			atomicOr(&dm[i], 1u << d);
		}
	}
	// voxel i pushed into t (its excess and residual sink link changed)
	static constexpr int kSinkBit = kConn <= 8 ? 6 : 30;
	__device__ static void touch(uint32_t *dm, int i) { set(dm, i, kSinkBit); }
	__device__ static uint32_t get(const uint32_t *dm, int i)
	{
		return kConn <= 8 ? (dm[i >> 2] >> (8 * (i & 3))) & 0xffu : dm[i];
	}
};

template <int kConn, int kTX, int kTY, int kTZ, typename RT>
__global__ __launch_bounds__(kTX *kTY, 6) void k_grid_discharge(GridView gv, const int32_t *__restrict__ list,
                                                             const int32_t *__restrict__ count, int class_stride,
                                                             int max_iters, uint8_t *__restrict__ tile_active)
{
	using T = Tile<kConn, kTX, kTY, kTZ>;
	using D = typename T::D;
	using Dirty = DirtyBits<kConn>;
	constexpr int K = T::K;
	const int tile = discharge_tile(list, count, class_stride, (int)blockIdx.x);
	if (tile < 0) {
		return;
	}
	int local;
	const GridView g = instance_view(gv, tile, local);
	int tx, ty, tz;
	tile_coords(g, local, tx, ty, tz);
	const int ox = tx * T::TX, oy = ty * T::TY, oz = tz * T::TZ;

	__shared__ int32_t hl[T::NP];
	__shared__ int32_t el[T::NV];
	__shared__ RT rl[K][T::NV];
	__shared__ uint32_t dm[Dirty::kWords(T::NV)];   // residual directions written during this visit
	// (step << 32 | level) of the sweep's steps, double-buffered by step parity: a wave that has
	// read step s's value and raced ahead writes step s + 1's into the other slot, so a slower wave
	// still reading step s's slot cannot see it; step s + 2 reuses the slot only after the barrier
	// of step s + 1, which every wave passes after its read of step s
	__shared__ unsigned long long key[2];
	__shared__ long long red[T::NT / 64];
	__shared__ uint8_t marks[27];

	const int tid = threadIdx.x;
	const int lx = tid % T::TX, ly = tid / T::TX;
	const int gx = ox + lx, gy = oy + ly;
	if (tid < 27) {
		marks[tid] = 0;
	}
	if (tid < 2) {
		key[tid] = 0;
	}
	for (int w = tid; w < Dirty::kWords(T::NV); w += T::NT) {
		dm[w] = 0;
	}
	// every global load of a group first, then its LDS stores: the halo labels with the excess,
	// then the residual planes in groups of kGroup directions
	const uint32_t xy = slice_xy(g, gx, gy), sxy = (uint32_t)(g.nx * g.ny);
	uint32_t valid = 0;  // bit k: voxel k of the column lies inside the volume
	int32_t rtr[T::TZ];
	{
		HaloLabels<T> halo;
		fetch_label_tile<T>(g, halo, ox, oy, oz);
		int32_t ev[T::TZ];
#pragma unroll
		for (int k = 0; k < T::TZ; ++k) {
			valid |= (gx < g.nx && gy < g.ny && oz + k < g.nz) ? (1u << k) : 0u;
			ev[k] = 0;
			rtr[k] = 0;
			if (valid & (1u << k)) {
				ev[k] = tile_at(g.e, g, oz, k, xy, sxy);
				rtr[k] = tile_at(g.rt, g, oz, k, xy, sxy);
			}
		}
		store_label_tile<T>(hl, halo);
#pragma unroll
		for (int k = 0; k < T::TZ; ++k) {
			el[T::vidx(lx, ly, k)] = ev[k];
		}
	}
	constexpr int kGroup = 6;
#pragma unroll
	for (int d0 = 0; d0 < K; d0 += kGroup) {
		int32_t rv[kGroup][T::TZ];
#pragma unroll
		for (int j = 0; j < kGroup; ++j) {
#pragma unroll
			for (int k = 0; k < T::TZ; ++k) {
				rv[j][k] = 0;
				if (d0 + j < K && (valid & (1u << k))) {
					rv[j][k] = tile_at(g.r + (int64_t)(d0 + j) * g.nvox, g, oz, k, xy, sxy);
				}
			}
		}
#pragma unroll
		for (int j = 0; j < kGroup; ++j) {
#pragma unroll
			for (int k = 0; k < T::TZ; ++k) {
				if (d0 + j < K) {
					rl[d0 + j][T::vidx(lx, ly, k)] = (RT)rv[j][k];
				}
			}
		}
	}
	// 6-connected: flow pushed across the tile faces during the visit. The z-column of thread
	// (lx, ly) touches at most one x face and one y face, and the two z faces at its ends.
	int32_t out_x[T::TZ], out_y[T::TZ], out_zlo = 0, out_zhi = 0;
#pragma unroll
	for (int k = 0; k < T::TZ; ++k) {
		out_x[k] = 0;
		out_y[k] = 0;
	}
	__syncthreads();

	unsigned int step = 0;
	long long sink = 0;
	for (int it = 0; it < max_iters; ++it) {
		// ---- region relabel: exact distances to t inside the tile, halo labels fixed
		uint32_t open[T::TZ];
#pragma unroll
		for (int k = 0; k < T::TZ; ++k) {
			const int i = T::vidx(lx, ly, k);
			uint32_t o = 0;
#pragma unroll
			for (int d = 0; d < K; ++d) {
				o |= rl[d][i] > 0 ? (1u << d) : 0u;
			}
			open[k] = o;
		}
		// (no barrier before these writes: the last reads of hl, the admissible-direction pass of
		// the previous iteration, are separated from them by the sweep's barriers)
#pragma unroll
		for (int k = 0; k < T::TZ; ++k) {
			if (valid & (1u << k)) {
				hl[T::pidx(lx, ly, k)] = rtr[k] > 0 ? 1 : kLabelInf;
			}
		}
		__syncthreads();
		relax_tile<T>(hl, open, valid, lx, ly);

		// ---- per own voxel, fixed for the push sweep below: its label, its admissible directions
		// (open, neighbour one level lower; a push never opens one, since it goes one level down)
		// and whether it held excess when the sweep began; and the top level holding excess
		int32_t lab[T::TZ];
		AdmissibleDirs<K, T::TZ> adm;
		uint32_t held = 0;
		int top = 0;
#pragma unroll
		for (int k = 0; k < T::TZ; ++k) {
			const int p = T::pidx(lx, ly, k);
			lab[k] = kLabelInf;
			if (!(valid & (1u << k))) {
				continue;
			}
			const int32_t l = hl[p];
			lab[k] = l;
			uint32_t a = 0;
#pragma unroll
			for (int d = 0; d < K; ++d) {
				if ((open[k] & (1u << d)) && hl[p + T::poff(d)] == l - 1) {
					a |= 1u << d;
				}
			}
			adm.set(k, a);
			if (el[T::vidx(lx, ly, k)] > 0 && l < kLabelInf) {
				held |= 1u << k;
				top = l > top ? l : top;
			}
		}
		++step;
		top = wave_max_i32(top);
		if ((tid & 63) == 0) {
			atomicMax(&key[step & 1], ((unsigned long long)step << 32) | (unsigned)top);
		}
		__syncthreads();
		int level = (int)(key[step & 1] & 0xffffffffull);
		if (level == 0) {
			break;
		}
		// ---- top-down push sweep over the levels that hold excess. The labels are exact BFS
		// distances inside the tile, so a voxel of level L pushes only to voxels of level L - 1:
		// the next level to visit is L - 1 if this step pushed inside the tile, else the highest
		// lower level whose voxels held excess when the sweep began (a voxel below the current
		// level has received nothing yet).
		while (level >= 1) {
			int next = 0;
#pragma unroll
			for (int k = 0; k < T::TZ; ++k) {
				if (lab[k] != level) {
					if (lab[k] < level && lab[k] > next && (held & (1u << k))) {
						next = lab[k];
					}
					continue;
				}
				const int i = T::vidx(lx, ly, k);
				int32_t ex = el[i];
				if (ex <= 0) {
					continue;
				}
				if (rtr[k] > 0) {
					const int32_t f = ex < rtr[k] ? ex : rtr[k];
					rtr[k] -= f;
					ex -= f;
					sink += f;
					Dirty::touch(dm, i);
				}
				const uint32_t a = adm.get(k);
				// sends f along direction d: the own residual, and the neighbour's excess and reverse
				// residual inside the tile, or the face accumulator (6C) beyond it
				auto send = [&](int d, int32_t rv, int32_t f) {
					rl[d][i] = (RT)(rv - f);
					Dirty::set(dm, i, d);
					const int nx = lx + D::dx(d), ny = ly + D::dy(d), nz = k + D::dz(d);
					if (nx >= 0 && nx < T::TX && ny >= 0 && ny < T::TY && nz >= 0 && nz < T::TZ) {
						const int j = i + T::voff(d);
						rl[opposite<kConn>(d)][j] = (RT)(rl[opposite<kConn>(d)][j] + f);
						Dirty::set(dm, j, opposite<kConn>(d));
						atomicAdd(&el[j], f);
						next = level - 1 > next ? level - 1 : next;
					} else if (kConn == 6) {
						if (d == 0 || d == 3) {
							out_x[k] += f;
						} else if (d == 1 || d == 4) {
							out_y[k] += f;
						} else if (d == 2) {
							out_zlo += f;
						} else {
							out_zhi += f;
						}
					}
				};
				if (kConn == 6) {
#pragma unroll
					for (int d = 0; d < K; ++d) {
						if (ex > 0 && (a & (1u << d))) {
							const int32_t rv = rl[d][i];
							if (rv > 0) {
								const int32_t f = ex < rv ? ex : rv;
								ex -= f;
								send(d, rv, f);
							}
						}
					}
				} else {
					// 26 directions: the admissible residuals of a group are read together and split
					// greedily in registers (the same split, direction by direction), so a voxel waits
					// for one LDS latency per group instead of one per direction
					constexpr int kGroup = 7;
#pragma unroll
					for (int d0 = 0; d0 < K; d0 += kGroup) {
						if (ex <= 0 || !(a & (((1u << kGroup) - 1u) << d0))) {
							continue;
						}
						int32_t rv[kGroup];
#pragma unroll
						for (int j = 0; j < kGroup; ++j) {
							const int d = d0 + j;
							rv[j] = (d < K && (a & (1u << d))) ? (int32_t)rl[d < K ? d : 0][i] : 0;
						}
#pragma unroll
						for (int j = 0; j < kGroup; ++j) {
							const int d = d0 + j;
							const int32_t f = ex < rv[j] ? ex : rv[j];
							if (d < K && f > 0) {
								ex -= f;
								send(d, rv[j], f);
							}
						}
					}
				}
				el[i] = ex;
			}
			++step;
			next = wave_max_i32(next);
			if ((tid & 63) == 0) {
				atomicMax(&key[step & 1], ((unsigned long long)step << 32) | (unsigned)next);
			}
			__syncthreads();
			level = (int)(key[step & 1] & 0xffffffffull);
		}
		// when the sweep left no excess in the tile, another region relabel would only find that
		// out; the labels in hl stay valid (pushes along admissible edges keep them valid) and are
		// what the write-back stores
		int left = 0;
#pragma unroll
		for (int k = 0; k < T::TZ; ++k) {
			left |= ((valid >> k) & 1u) && el[T::vidx(lx, ly, k)] > 0;
		}
		if (!__syncthreads_or(left)) {
			break;
		}
	}

	// ---- write back what changed; pushes that left the tile go to the neighbouring voxels
	int active = 0;
#pragma unroll
	for (int k = 0; k < T::TZ; ++k) {
		if (!(valid & (1u << k))) {
			continue;
		}
		const int p = T::pidx(lx, ly, k);
		const int i = T::vidx(lx, ly, k);
		const int32_t ev = el[i], lab = hl[p];
		active |= ev > 0 && lab < kLabelInf;
		tile_at(g.h, g, oz, k, xy, sxy) = lab;  // (a store is cheaper than keeping the loaded labels in registers)
		// every change of e, rt or r[d] set a bit of the voxel's dirty byte / word
		const uint32_t dirty = Dirty::get(dm, i);
		if (!dirty) {
			continue;
		}
		tile_at(g.e, g, oz, k, xy, sxy) = ev;
		tile_at(g.rt, g, oz, k, xy, sxy) = rtr[k];
		uint32_t st = (ev > 0 ? kStateExcess : 0u) | (rtr[k] > 0 ? kStateSink : 0u);
#pragma unroll
		for (int d = 0; d < K; ++d) {
			st |= rl[d][i] > 0 ? (1u << d) : 0u;
		}
		tile_at(g.omask, g, oz, k, xy, sxy) = st;
#pragma unroll
		for (int d = 0; d < K; ++d) {
			if (!(dirty & (1u << d))) {
				continue;
			}
			const int32_t nr = rl[d][i];
			int32_t *plane = g.r + (int64_t)d * g.nvox;
			if (kConn != 6) {
				const int slot = leaving_slot<T>(lx, ly, k, d);
				if (slot >= 0) {
					// only this voxel pushes along this edge during the phase: the drop is the flow sent
					const int32_t f = tile_at(plane, g, oz, k, xy, sxy) - nr;
					if (f > 0) {
						const int uk = k + D::dz(d);
						const uint32_t uxy = slice_xy(g, gx + D::dx(d), gy + D::dy(d));
						atomicAdd(&tile_at(g.e, g, oz, uk, uxy, sxy), f);
						atomicAdd(&tile_at(g.r + (int64_t)opposite<kConn>(d) * g.nvox, g, oz, uk, uxy, sxy), f);
						atomicOr(&tile_at(g.omask, g, oz, uk, uxy, sxy), kStateExcess | (1u << opposite<kConn>(d)));
						marks[slot] = 1;
					}
				}
			}
			tile_at(plane, g, oz, k, xy, sxy) = nr;
		}
	}
	if (kConn == 6) {
		// flow that crossed the faces goes to the neighbouring voxels (atomics without return)
		auto send = [&](int k, int d, int32_t f) {
			if (f <= 0) {
				return;
			}
			const int uk = k + D::dz(d);
			const uint32_t uxy = slice_xy(g, gx + D::dx(d), gy + D::dy(d));
			atomicAdd(&tile_at(g.e, g, oz, uk, uxy, sxy), f);
			atomicAdd(&tile_at(g.r + (int64_t)opposite<kConn>(d) * g.nvox, g, oz, uk, uxy, sxy), f);
			atomicOr(&tile_at(g.omask, g, oz, uk, uxy, sxy), kStateExcess | (1u << opposite<kConn>(d)));
			marks[leaving_slot<T>(lx, ly, k, d)] = 1;
		};
#pragma unroll
		for (int k = 0; k < T::TZ; ++k) {
			send(k, lx == 0 ? 0 : 3, out_x[k]);
			send(k, ly == 0 ? 1 : 4, out_y[k]);
		}
		send(0, 2, out_zlo);
		send(T::TZ - 1, 5, out_zhi);
	}
	const long long total = block_sum_i64<T::NT>(sink, red);
	active = __syncthreads_or(active);
	if (tid == 0) {
		if (total) {
			atomicAdd(flow_slot(g, tile), (unsigned long long)total);
		}
		tile_active[tile] = active ? 1 : 0;
		atomicAdd(counter_slot(g, 0, tile), 1ull);
	}
	if (tid < 27 && marks[tid]) {
		mark_neighbour_tile(g, tile_active, tx, ty, tz, tid);
	}
}

// ---------------------------------------------------------------------------------------------
// S_min closure pass: s(u) = 1 if u has excess or some v with s(v) = 1 has residual v -> u > 0.
// s starts at 0 (a fill before the first pass) and bits only rise, so concurrently running tiles
// may set bits in each other's voxels. first_pass seeds from the excess; the open directions come
// from the state word. With reached set, a pass that puts a voxel with a residual sink link into the
// closure stores 1 there: some excess can still reach t, the preflow is not maximum yet.
template <int kConn, int kTX, int kTY, int kTZ>
__global__ __launch_bounds__(kTX *kTY) void k_grid_source_closure(GridView gv, const int32_t *__restrict__ list,
                                                                  const int32_t *__restrict__ count, int first_pass,
                                                                  uint8_t *__restrict__ dirty_next,
                                                                  int32_t *__restrict__ reached,
                                                                  const SolveControl *ctl)
{
	using T = Tile<kConn, kTX, kTY, kTZ>;
	using D = typename T::D;
	const int tile = list[blockIdx.x];  // in bounds: the grid never exceeds the list's capacity
	if ((int)blockIdx.x >= *count || !gate_open(ctl, kGateClosure | kGateUnlessReached)) {
		return;
	}
	int local;
	const GridView g = instance_view(gv, tile, local);
	int tx, ty, tz;
	tile_coords(g, local, tx, ty, tz);
	const int ox = tx * T::TX, oy = ty * T::TY, oz = tz * T::TZ;

	__shared__ uint8_t sl[T::NV];
	__shared__ uint8_t marks[27];
	const int tid = threadIdx.x;
	const int lx = tid % T::TX, ly = tid / T::TX;
	const int gx = ox + lx, gy = oy + ly;
	if (tid < 27) {
		marks[tid] = 0;
	}
	const uint32_t xy = slice_xy(g, gx, gy), sxy = (uint32_t)(g.nx * g.ny);
	bool valid[T::TZ];
	uint32_t open[T::TZ];
	uint8_t before[T::TZ];
	uint32_t sink = 0;  // bit k: own voxel k has a residual sink link
	// all loads first, then the LDS stores
#pragma unroll
	for (int k = 0; k < T::TZ; ++k) {
		valid[k] = gx < g.nx && gy < g.ny && oz + k < g.nz;
		before[k] = 0;
		open[k] = 0;
		if (valid[k]) {
			before[k] = tile_at(g.s, g, oz, k, xy, sxy);
			open[k] = tile_at(g.omask, g, oz, k, xy, sxy);
		}
	}
#pragma unroll
	for (int k = 0; k < T::TZ; ++k) {
		const uint32_t st = open[k];
		sl[T::vidx(lx, ly, k)] = ((first_pass && (st & kStateExcess)) || before[k]) ? 1 : 0;
		open[k] = st & ((1u << T::K) - 1u);
		sink |= (st & kStateSink) ? (1u << k) : 0u;
	}
	__syncthreads();
	for (;;) {
		int changed = 0;
#pragma unroll
		for (int k = 0; k < T::TZ; ++k) {
			const int i = T::vidx(lx, ly, k);
			if (!valid[k] || !sl[i]) {
				continue;
			}
#pragma unroll
			for (int d = 0; d < T::K; ++d) {
				if (!(open[k] & (1u << d))) {
					continue;
				}
				const int nx = lx + D::dx(d), ny = ly + D::dy(d), nz = k + D::dz(d);
				if (nx >= 0 && nx < T::TX && ny >= 0 && ny < T::TY && nz >= 0 && nz < T::TZ) {
					const int j = i + T::voff(d);
					if (!sl[j]) {
						sl[j] = 1;
						changed = 1;
					}
				}
			}
		}
		if (!__syncthreads_or(changed)) {
			break;
		}
	}
	// write back own voxels; push the closure across the tile faces
	bool hit = false;
#pragma unroll
	for (int k = 0; k < T::TZ; ++k) {
		const int i = T::vidx(lx, ly, k);
		if (!valid[k]) {
			continue;
		}
		hit |= sl[i] && ((sink >> k) & 1u);
		if (sl[i] != before[k]) {
			tile_at(g.s, g, oz, k, xy, sxy) = sl[i];
		}
		if (!sl[i]) {
			continue;
		}
#pragma unroll
		for (int d = 0; d < T::K; ++d) {
			if (!(open[k] & (1u << d))) {
				continue;
			}
			const int slot = leaving_slot<T>(lx, ly, k, d);
			if (slot < 0) {
				continue;
			}
			uint8_t &su = tile_at(g.s, g, oz, k + D::dz(d), slice_xy(g, gx + D::dx(d), gy + D::dy(d)), sxy);
			if (!su) {
				su = 1;
				marks[slot] = 1;
			}
		}
	}
	if (reached && __ballot(hit) && (tid & 63) == 0) {
		*reached = 1;  // every writer stores the same value
	}
	__syncthreads();
	if (tid < 27 && marks[tid]) {
		mark_neighbour_tile(g, dirty_next, tx, ty, tz, tid);
	}
	if (tid == 0) {
		atomicAdd(counter_slot(g, 2, tile), 1ull);
	}
}

// ---------------------------------------------------------------------------------------------
// total[i] = sum of the kSlots slots of entry i (an instance's flow, or one visit counter).
__global__ __launch_bounds__(256) void k_sum_slots(const unsigned long long *__restrict__ slots, int64_t n,
                                                   unsigned long long *__restrict__ total)
{
	const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
	if (i >= n) {
		return;
	}
	unsigned long long sum = 0;
#pragma unroll
	for (int j = 0; j < kSlots; ++j) {
		sum += slots[i * kSlotWords + j * kSlotStride];
	}
	total[i] = sum;
}

// The same compaction by one workgroup, for up to kCompactBlockTiles tiles: thread t takes the
// chunk [t·32m, (t+1)·32m) (m = ceil(ntiles / 32768)) in pieces of 32 tiles, each read with two
// 16-byte loads per array (the flag arrays are padded to whole 32-byte pieces); a workgroup scan of
// the chunk counts places the chunks in order, so the list comes out sorted and the count is
// written rather than accumulated (no fill beforehand). With flags2 set, a second flag array (no
// colour filter, no list) is counted into count2 in the same pass. With one piece a thread, a wave
// with more than kDirect list entries stores them through an LDS stage by consecutive lanes, and
// consumed flags go back as whole 16-byte words.
constexpr int kCompactThreads = 1024;
// one piece per thread for up to 32768 tiles (the solver uses it with one piece; beyond, lists go
// through the multi-workgroup passes: k_compact_tiles, or k_order_regions for the ordered
// discharge lists, which measured faster than eight pieces a thread in one workgroup)
constexpr int kCompactBlockTiles = 32 * kCompactThreads;

__device__ inline int block_exclusive_scan_1024(int v, int *wave_sums, int &total)
{
	const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
	int inc = v;
#pragma unroll
	for (int o = 1; o < 64; o <<= 1) {
		const int t = __shfl_up(inc, (unsigned)o, 64);
		if (lane >= o) {
			inc += t;
		}
	}
	if (lane == 63) {
		wave_sums[wave] = inc;
	}
	__syncthreads();
	int before = 0;
	total = 0;
#pragma unroll
	for (int w = 0; w < kCompactThreads / 64; ++w) {
		const int s = wave_sums[w];
		before += w < wave ? s : 0;
		total += s;
	}
	return before + inc - v;
}

// ---------------------------------------------------------------------------------------------
// Compact the indices of flagged tiles (optionally of one colour) into a list, for any number of
// tiles: workgroups of 1024 threads take 4096 tiles (4 per thread, one 4-byte load per array; the
// flag arrays are padded to whole 32-byte pieces), scan their selections and reserve their part
// of the list with one atomic each (per-wave atomics on the one count serialise in the L2). The
// count must be zero beforehand; zero_next (if set) is cleared, so a solver that alternates
// between two count words needs no fill before a compaction. colours[t] is the colour of tile t
// inside its instance. With clear set, the selected flags are reset (the pass that takes the list
// consumes them); with list null, only the count is taken.
constexpr int kCompactTilesPerThread = 4;
constexpr int kCompactTilesPerGroup = kCompactTilesPerThread * kCompactThreads;

__global__ __launch_bounds__(kCompactThreads) void k_compact_tiles(uint8_t *__restrict__ flags, int ntiles,
                                                                   const uint8_t *__restrict__ colours, int colour,
                                                                   int clear, int32_t *__restrict__ list,
                                                                   int32_t *__restrict__ count,
                                                                   int32_t *__restrict__ zero_next,
                                                                   const SolveControl *ctl, unsigned gate)
{
	__shared__ int sums[kCompactThreads / 64];
	__shared__ int base;
	if (zero_next && blockIdx.x == 0 && threadIdx.x == 0) {
		*zero_next = 0;  // the count of the next compaction (it alternates between two words)
	}
	if (!gate_open(ctl, gate)) {
		return;  // the count stays at zero
	}
	const int t0 = ((int)blockIdx.x * kCompactThreads + (int)threadIdx.x) * kCompactTilesPerThread;
	uint32_t sel = 0;
	if (t0 < ntiles) {
		const uint32_t f = *reinterpret_cast<const uint32_t *>(flags + t0);
		const uint32_t c = colour >= 0 ? *reinterpret_cast<const uint32_t *>(colours + t0) : 0u;
#pragma unroll
		for (int j = 0; j < kCompactTilesPerThread; ++j) {
			const bool on = ((f >> (8 * j)) & 0xffu) != 0u &&
			                (colour < 0 || ((c >> (8 * j)) & 0xffu) == (uint32_t)colour) && t0 + j < ntiles;
			sel |= (on ? 1u : 0u) << j;
		}
	}
	int total = 0;
	int pos = block_exclusive_scan_1024(__popc(sel), sums, total);
	if (threadIdx.x == 0) {
		base = total ? atomicAdd(count, total) : 0;
	}
	__syncthreads();
	pos += base;
	for (uint32_t m = sel; m; m &= m - 1) {
		const int i = t0 + __ffs(m) - 1;
		if (list) {
			list[pos++] = i;
		}
		if (clear) {
			flags[i] = 0;
		}
	}
}

// bit j set where byte j of the 16 bytes q is nonzero (nonzero) or equals colour
__device__ inline uint32_t byte_mask16(uint4 q, int colour, bool nonzero)
{
	const uint32_t w[4] = {q.x, q.y, q.z, q.w};
	uint32_t m = 0;
#pragma unroll
	for (int j = 0; j < 16; ++j) {
		const uint32_t b = (w[j >> 2] >> (8 * (j & 3))) & 0xffu;
		m |= ((nonzero ? b != 0u : b == (uint32_t)colour) ? 1u : 0u) << j;
	}
	return m;
}

// selection masks of the 32 tiles from lo (lo < ntiles, a multiple of 32)
// (and with kKeep the flag bytes in keep0, keep1)
template <bool kKeep>
__device__ inline void compact_piece(const uint8_t *flags, const uint8_t *colours, int colour, const uint8_t *flags2,
                                     int lo, int ntiles, uint32_t &sel, uint32_t &sel2, uint4 &keep0, uint4 &keep1)
{
	// all loads first: one memory latency for the three arrays
	const uint4 *f = reinterpret_cast<const uint4 *>(flags + lo);
	const uint4 f0 = f[0], f1 = f[1];
	if constexpr (kKeep) {
		keep0 = f0;
		keep1 = f1;
	}
	uint4 c0 = {}, c1 = {}, g0 = {}, g1 = {};
	if (colour >= 0) {
		const uint4 *c = reinterpret_cast<const uint4 *>(colours + lo);
		c0 = c[0];
		c1 = c[1];
	}
	if (flags2) {
		const uint4 *f2 = reinterpret_cast<const uint4 *>(flags2 + lo);
		g0 = f2[0];
		g1 = f2[1];
	}
	sel = byte_mask16(f0, 0, true) | (byte_mask16(f1, 0, true) << 16);
	if (colour >= 0) {
		sel &= byte_mask16(c0, colour, false) | (byte_mask16(c1, colour, false) << 16);
	}
	sel2 = flags2 ? byte_mask16(g0, 0, true) | (byte_mask16(g1, 0, true) << 16) : 0u;
	const int n = ntiles - lo;
	if (n < 32) {
		sel &= (1u << n) - 1u;
		sel2 &= (1u << n) - 1u;
	}
}

// With work set (a per-tile estimate) and more than kOrderMin tiles listed (fewer all run at once),
// the list is ordered by decreasing work class (kWorkClasses classes of work / 4, the last one
// open-ended), in tile order within a class: the heaviest tiles start first and do not finish
// after everyone else. The per-class counts travel packed, 16 bits a class in two 64-bit words, so
// one workgroup scan orders all classes.
constexpr int kOrderMin = 2048;

__device__ inline int work_class(int w)
{
	const int b = w >> 2;
	return kWorkClasses - 1 - (b < kWorkClasses - 1 ? b : kWorkClasses - 1);
}

// Masks of the 32 tiles of a piece by work class (bit j of cm[c]: tile j is of class c), from
// two 16-byte loads of their work bytes (one dependent load per listed tile would put a memory
// latency on every step of the loops over the selection bits; the masks need no register array
// indexed at run time).
__device__ inline void class_masks(const uint8_t *work, int first, uint32_t (&cm)[kWorkClasses])
{
	const uint4 *p = reinterpret_cast<const uint4 *>(work + first);
	const uint4 lo = p[0], hi = p[1];
	const uint32_t w[8] = {lo.x, lo.y, lo.z, lo.w, hi.x, hi.y, hi.z, hi.w};
#pragma unroll
	for (int c = 0; c < kWorkClasses; ++c) {
		cm[c] = 0;
	}
#pragma unroll
	for (int j = 0; j < 32; ++j) {
		const int cls = work_class((int)((w[j >> 2] >> (8 * (j & 3))) & 0xffu));
#pragma unroll
		for (int c = 0; c < kWorkClasses; ++c) {
			cm[c] |= (cls == c ? 1u : 0u) << j;
		}
	}
}

// The work bytes of a piece for the variant with several pieces a thread (per-class masks of every
// piece would not fit its registers): the class of one tile at a time.
struct PieceWork {
	uint4 lo, hi;
	__device__ PieceWork(const uint8_t *work, int first)
	{
		const uint4 *p = reinterpret_cast<const uint4 *>(work + first);
		lo = p[0];
		hi = p[1];
	}
	// work class of tile j (0 <= j < 32) of the piece
	__device__ int cls(int j) const
	{
		const int d = j >> 2;
		const uint32_t a = (d & 1) ? lo.y : lo.x, b = (d & 1) ? lo.w : lo.z;
		const uint32_t c = (d & 1) ? hi.y : hi.x, e = (d & 1) ? hi.w : hi.z;
		const uint32_t w = d < 4 ? ((d & 2) ? b : a) : ((d & 2) ? e : c);
		return work_class((int)((w >> (8 * (j & 3))) & 0xffu));
	}
};

// Exclusive scan of (lo, hi) pairs of packed counts over the workgroup; totals in tlo, thi.
template <int kThreads = kCompactThreads>
__device__ inline void block_exclusive_scan_pair(unsigned long long &lo, unsigned long long &hi,
                                                 unsigned long long (*wave_sums)[2], unsigned long long &tlo,
                                                 unsigned long long &thi)
{
	const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
	unsigned long long a = lo, b = hi;
#pragma unroll
	for (int o = 1; o < 64; o <<= 1) {
		const unsigned long long ta = __shfl_up(a, (unsigned)o, 64), tb = __shfl_up(b, (unsigned)o, 64);
		if (lane >= o) {
			a += ta;
			b += tb;
		}
	}
	if (lane == 63) {
		wave_sums[wave][0] = a;
		wave_sums[wave][1] = b;
	}
	__syncthreads();
	unsigned long long before_a = 0, before_b = 0;
	tlo = thi = 0;
#pragma unroll
	for (int w = 0; w < kThreads / 64; ++w) {
		const unsigned long long sa = wave_sums[w][0], sb = wave_sums[w][1];
		before_a += w < wave ? sa : 0;
		before_b += w < wave ? sb : 0;
		tlo += sa;
		thi += sb;
	}
	lo = before_a + a - lo;
	hi = before_b + b - hi;
}

// q with byte j zeroed where bit j of bits is set
__device__ inline uint4 zero_bytes16(uint4 q, uint32_t bits)
{
	uint32_t w[4] = {q.x, q.y, q.z, q.w};
#pragma unroll
	for (int i = 0; i < 4; ++i) {
		uint32_t keep = 0;
#pragma unroll
		for (int j = 0; j < 4; ++j) {
			keep |= ((bits >> (4 * i + j)) & 1u) ? 0u : 0xffu << (8 * j);
		}
		w[i] &= keep;
	}
	return make_uint4(w[0], w[1], w[2], w[3]);
}

// LDS written by lanes of this wave is read by other lanes of it after this point (a wave's LDS
// operations complete in order; the fences keep the compiler from moving them across)
__device__ inline void wave_lds_sync()
{
	__builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
	__builtin_amdgcn_wave_barrier();
	__builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// list entries staged per wave and pass, for waves with more than kDirect entries
constexpr int kStage = 512;
constexpr int kDirect = 256;

template <int kCompactMaxPieces>
__global__ __launch_bounds__(kCompactThreads) void k_compact_block(uint8_t *__restrict__ flags, int ntiles,
                                                                   const uint8_t *__restrict__ colours, int colour,
                                                                   int clear, int32_t *__restrict__ list,
                                                                   int32_t *__restrict__ count,
                                                                   const uint8_t *__restrict__ flags2,
                                                                   int32_t *__restrict__ count2,
                                                                   const uint8_t *__restrict__ work,
                                                                   const SolveControl *ctl, unsigned gate)
{
	__shared__ int sums[2][kCompactThreads / 64];
	__shared__ unsigned long long class_sums[kCompactThreads / 64][2];
	if (!gate_open(ctl, gate)) {
		// a closed gate lists nothing
		if (threadIdx.x == 0) {
			*count = 0;
			if (flags2) {
				*count2 = 0;
			}
		}
		return;
	}
	const int pieces = (ntiles + 32 * kCompactThreads - 1) / (32 * kCompactThreads);  // <= kCompactMaxPieces
	const int base = (int)threadIdx.x * 32 * pieces;
	uint32_t sel[kCompactMaxPieces];
	// one piece keeps its flag bytes for the clear
	constexpr bool kHold = kCompactMaxPieces == 1;
	uint4 held0 = {}, held1 = {};
	int n = 0, n2 = 0;
#pragma unroll
	for (int q = 0; q < kCompactMaxPieces; ++q) {
		sel[q] = 0;
		const int lo = base + 32 * q;
		if (q < pieces && lo < ntiles) {
			uint32_t s2;
			compact_piece<kHold>(flags, colours, colour, flags2, lo, ntiles, sel[q], s2, held0, held1);
			n += __popc(sel[q]);
			n2 += __popc(s2);
		}

	}
	int total = 0;
	int pos = block_exclusive_scan_1024(n, sums[0], total);
	if (threadIdx.x == 0) {
		*count = total;
	}
	if (flags2) {
		int total2 = 0;
		block_exclusive_scan_1024(n2, sums[1], total2);
		if (threadIdx.x == 0) {
			*count2 = total2;
		}
	}
	// uniform; the packed class counts hold 16 bits, so longer lists keep tile order
	const bool ordered = list && work && total > kOrderMin && total <= 0xffff;
	const int lane = (int)threadIdx.x & 63, wave = (int)threadIdx.x >> 6;
	if (clear) {
		if constexpr (kHold) {
			// the piece written back whole, the selected bytes zeroed
			if (sel[0]) {
				uint4 *f = reinterpret_cast<uint4 *>(flags + base);
				f[0] = zero_bytes16(held0, sel[0] & 0xffffu);
				f[1] = zero_bytes16(held1, sel[0] >> 16);
			}
		} else {
#pragma unroll
			for (int q = 0; q < kCompactMaxPieces; ++q) {
				for (uint32_t m = sel[q]; m; m &= m - 1) {
					flags[base + 32 * q + __ffs(m) - 1] = 0;
				}
			}
		}
	}
	if (!list) {
		return;
	}
	if (!ordered) {
		const int first = __shfl(pos, 0, 64), mine = pos - first, wn = sums[0][wave];
		// (staged only with one piece a thread: with more, a thread's entries span many stage
		// passes and it would walk its selection bits in every pass)
		if (!kHold || wn <= kDirect) {
#pragma unroll
			for (int q = 0; q < kCompactMaxPieces; ++q) {
				for (uint32_t m = sel[q]; m; m &= m - 1) {
					list[pos++] = base + 32 * q + __ffs(m) - 1;
				}
			}
			return;
		}
		// A wave with many entries places them in its stage in LDS, kStage at a time, and copies
		// them out by consecutive lanes (a thread writing its own entries straight to the list
		// puts every lane of a store on another cache line).
		__shared__ int stage[kCompactThreads / 64][kStage];
		int *st = stage[wave];
		for (int c0 = 0; c0 < wn; c0 += kStage) {
			if (n && mine < c0 + kStage && mine + n > c0) {
				int k = mine;
#pragma unroll
				for (int q = 0; q < kCompactMaxPieces; ++q) {
					for (uint32_t m = sel[q]; m; m &= m - 1, ++k) {
						if (k >= c0 && k < c0 + kStage) {
							st[k - c0] = base + 32 * q + __ffs(m) - 1;
						}
					}
				}
			}
			wave_lds_sync();
			const int cnt = wn - c0 < kStage ? wn - c0 : kStage;
			for (int e = lane; e < cnt; e += 64) {
				list[first + c0 + e] = st[e];
			}
			wave_lds_sync();
		}
		return;
	}
	// per-class counts, 16 bits each: classes 0-3 in lo, 4-7 in hi
	unsigned long long lo = 0, hi = 0;
#pragma unroll
	for (int q = 0; q < kCompactMaxPieces; ++q) {
		if (!sel[q]) {
			continue;
		}
		if constexpr (kHold) {
			uint32_t cm[kWorkClasses];
			class_masks(work, base + 32 * q, cm);
#pragma unroll
			for (int c = 0; c < kWorkClasses; ++c) {
				const unsigned long long k = (unsigned long long)__popc(cm[c] & sel[q]) << (16 * (c & 3));
				lo += c < 4 ? k : 0ull;
				hi += c < 4 ? 0ull : k;
			}
		} else {
			const PieceWork pw(work, base + 32 * q);
			for (uint32_t m = sel[q]; m; m &= m - 1) {
				const int c = pw.cls(__ffs(m) - 1);
				const unsigned long long one = 1ull << (16 * (c & 3));
				lo += c < 4 ? one : 0ull;
				hi += c < 4 ? 0ull : one;
			}
		}
	}
	unsigned long long tlo = 0, thi = 0;
	block_exclusive_scan_pair(lo, hi, class_sums, tlo, thi);
	// class c starts after the totals of the classes before it
	int offset[kWorkClasses];
	int start = 0;
#pragma unroll
	for (int c = 0; c < kWorkClasses; ++c) {
		const unsigned long long t = c < 4 ? tlo : thi, e = c < 4 ? lo : hi;
		offset[c] = start + (int)((e >> (16 * (c & 3))) & 0xffffull);
		start += (int)((t >> (16 * (c & 3))) & 0xffffull);
	}
#pragma unroll
	for (int q = 0; q < kCompactMaxPieces; ++q) {
		if (!sel[q]) {
			continue;
		}
		if constexpr (kHold) {
			uint32_t cm[kWorkClasses];
			class_masks(work, base + 32 * q, cm);
#pragma unroll
			for (int c = 0; c < kWorkClasses; ++c) {
				for (uint32_t m = cm[c] & sel[q]; m; m &= m - 1) {
					list[offset[c]++] = base + 32 * q + __ffs(m) - 1;
				}
			}
		} else {
			const PieceWork pw(work, base + 32 * q);
			for (uint32_t m = sel[q]; m; m &= m - 1) {
				const int c = pw.cls(__ffs(m) - 1);
#pragma unroll
				for (int cc = 0; cc < kWorkClasses; ++cc) {
					if (cc == c) {
						list[offset[cc]++] = base + 32 * q + __ffs(m) - 1;
					}
				}
			}
		}
	}
}

// ---------------------------------------------------------------------------------------------
// Ordered compaction over many workgroups in one pass, for lists longer than one workgroup
// compacts quickly: the list is laid out in class regions (class c's tiles at list + c * stride,
// stride = the largest colour class, their number in cursors[c]), which the discharge walks in
// class order (discharge_tile), so no pass needs the totals of the classes before it. Workgroups
// take 1024 tiles (4 a thread, one 4-byte load per array), scan their per-class counts, and
// reserve their run in each class's region with one atomic per class (the eight issued together by
// eight lanes). The cursors must be zero beforehand; zero_next, the other set of a solver that
// alternates between two, is cleared. Within a class the order between workgroups follows their
// reservations, which does not matter: any order of the tiles of a phase gives the same phase
// (tiles of one colour never touch).
// selection (bit j) and packed per-class counts of the thread's 4 tiles from t0
__device__ inline uint32_t order_select(const uint8_t *flags, int ntiles, const uint8_t *colours, int colour,
                                        const uint8_t *work, int t0, uint32_t &cls4, unsigned long long &lo,
                                        unsigned long long &hi)
{
	uint32_t sel = 0;
	cls4 = 0;
	lo = hi = 0;
	if (t0 >= ntiles) {
		return 0;
	}
	const uint32_t f = *reinterpret_cast<const uint32_t *>(flags + t0);
	const uint32_t c = colour >= 0 ? *reinterpret_cast<const uint32_t *>(colours + t0) : 0u;
	const uint32_t w = *reinterpret_cast<const uint32_t *>(work + t0);
#pragma unroll
	for (int j = 0; j < kCompactTilesPerThread; ++j) {
		const bool on = ((f >> (8 * j)) & 0xffu) != 0u &&
		                (colour < 0 || ((c >> (8 * j)) & 0xffu) == (uint32_t)colour) && t0 + j < ntiles;
		const int k = work_class((int)((w >> (8 * j)) & 0xffu));
		sel |= (on ? 1u : 0u) << j;
		cls4 |= (uint32_t)k << (8 * j);
		if (on) {
			const unsigned long long one = 1ull << (16 * (k & 3));
			lo += k < 4 ? one : 0ull;
			hi += k < 4 ? 0ull : one;
		}
	}
	return sel;
}

constexpr int kOrderThreads = 256;  // 1024 tiles a workgroup: short scans, many workgroups at once

__global__ __launch_bounds__(kOrderThreads) void k_order_regions(const uint8_t *__restrict__ flags, int ntiles,
                                                                   const uint8_t *__restrict__ colours, int colour,
                                                                   const uint8_t *__restrict__ work,
                                                                   int32_t *__restrict__ list, int stride,
                                                                   int32_t *__restrict__ cursors,
                                                                   int32_t *__restrict__ zero_next,
                                                                   const SolveControl *ctl, unsigned gate)
{
	__shared__ unsigned long long sums[kOrderThreads / 64][2];
	__shared__ int base[kWorkClasses];
	if (blockIdx.x == 0 && threadIdx.x < kWorkClasses) {
		zero_next[threadIdx.x] = 0;
	}
	if (!gate_open(ctl, gate)) {
		return;  // the cursors stay at zero: an empty list
	}
	const int t0 = ((int)blockIdx.x * kOrderThreads + (int)threadIdx.x) * kCompactTilesPerThread;
	uint32_t cls4;
	unsigned long long lo, hi;
	const uint32_t sel = order_select(flags, ntiles, colours, colour, work, t0, cls4, lo, hi);
	unsigned long long tlo = 0, thi = 0;
	block_exclusive_scan_pair<kOrderThreads>(lo, hi, sums, tlo, thi);  // lo, hi: this thread's offsets per class
	if (threadIdx.x < kWorkClasses) {
		const int c = (int)threadIdx.x;
		const int n = (int)(((c < 4 ? tlo : thi) >> (16 * (c & 3))) & 0xffffull);
		base[c] = c * stride + (n ? atomicAdd(&cursors[c], n) : 0);
	}
	__syncthreads();
	int next[kWorkClasses];
#pragma unroll
	for (int c = 0; c < kWorkClasses; ++c) {
		next[c] = base[c] + (int)(((c < 4 ? lo : hi) >> (16 * (c & 3))) & 0xffffull);
	}
#pragma unroll
	for (int j = 0; j < kCompactTilesPerThread; ++j) {
		if (!((sel >> j) & 1u)) {
			continue;
		}
		const int k = (int)((cls4 >> (8 * j)) & 0xffu);
#pragma unroll
		for (int c = 0; c < kWorkClasses; ++c) {
			if (c == k) {
				list[next[c]++] = t0 + j;
			}
		}
	}
}

// ---------------------------------------------------------------------------------------------
// Controllers of the device-driven solve (single workgroups; SolveControl in grid_control.h).

// Sets the stage (one thread): the host continues a solve that stopped for kModeNeedWide.
__global__ void k_ctl_set_mode(SolveControl *ctl, int32_t mode)
{
	if (threadIdx.x == 0 && blockIdx.x == 0) {
		ctl->mode = mode;
	}
}

// After the init kernel: input errors fail the solve; residuals wider than the graph's discharge
// variant stop it for the host to switch variants.
__global__ void k_ctl_after_init(SolveControl *ctl, const int32_t *status, int wide_graph)
{
	if (threadIdx.x != 0 || blockIdx.x != 0 || ctl->mode != kModeInit) {
		return;
	}
	const int32_t s = *status;
	ctl->status = s;
	if (s & 3) {
		ctl->mode = kModeFailed;
	} else if ((s & 4) && !wide_graph) {
		ctl->mode = kModeNeedWide;
	} else {
		ctl->mode = kModeRelabelNew;
	}
}

// Start of a colour-sweep stage: the first sweep visits every tile of both colours when the stage
// starts from scratch (a mode in fresh_mask), none when it continues. sizes: the two colours' tile
// counts.
__global__ void k_ctl_stage_begin(SolveControl *ctl, unsigned fresh_mask, const int32_t *sizes)
{
	if (threadIdx.x != 0 || blockIdx.x != 0) {
		return;
	}
	const bool fresh = (fresh_mask >> ctl->mode) & 1u;
	ctl->first[0] = fresh ? sizes[0] : 0;
	ctl->first[1] = fresh ? sizes[1] : 0;
	if (fresh) {
		ctl->reached = 0;
	}
}

constexpr int kCtlThreads = 1024;

// Number of nonzero bytes of flags[0, n) (padded to whole 16-byte words) over one workgroup, in
// thread 0.
__device__ inline int ctl_count_flags(const uint8_t *flags, int n, int *wave_sums)
{
	int c = 0;
	const int words = (n + 15) / 16;
	for (int w = (int)threadIdx.x; w < words; w += kCtlThreads) {
		const uint4 q = reinterpret_cast<const uint4 *>(flags)[w];
		const uint32_t v[4] = {q.x, q.y, q.z, q.w};
#pragma unroll
		for (int j = 0; j < 16; ++j) {
			const int i = w * 16 + j;
			c += (i < n && ((v[j >> 2] >> (8 * (j & 3))) & 0xffu)) ? 1 : 0;
		}
	}
#pragma unroll
	for (int off = 32; off > 0; off >>= 1) {
		c += __shfl_xor(c, off, 64);
	}
	const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
	if (lane == 0) {
		wave_sums[wave] = c;
	}
	__syncthreads();
	int total = 0;
	if (threadIdx.x == 0) {
		for (int w = 0; w < kCtlThreads / 64; ++w) {
			total += wave_sums[w];
		}
	}
	__syncthreads();
	return total;
}

// End of the relabel stage: dirty tiles left continue it; otherwise a relabel that found no active
// tile ends the discharge rounds (the S_min closure follows), and one that did starts the
// discharge phases unless the round cap is reached.
__global__ __launch_bounds__(kCtlThreads) void k_ctl_relabel_end(SolveControl *ctl, const uint8_t *dirty,
                                                                 const uint8_t *active, int ntiles)
{
	__shared__ int sums[kCtlThreads / 64];
	const int m = *(volatile const int32_t *)&ctl->mode;
	if (m != kModeRelabelNew && m != kModeRelabelCont) {
		return;
	}
	const int nd = ctl_count_flags(dirty, ntiles, sums);
	const int na = ctl_count_flags(active, ntiles, sums);
	if (threadIdx.x == 0) {
		ctl->dirty = nd;
		if (nd > 0) {
			ctl->mode = kModeRelabelCont;
			return;
		}
		ctl->active = na;
		const int rounds = ctl->rounds + 1;
		ctl->rounds = rounds;
		if (na == 0) {
			ctl->mode = kModeFinalNew;
		} else if (rounds - 1 >= ctl->max_rounds) {
			ctl->mode = kModeNotConverged;
		} else {
			ctl->mode = kModeDischarge;
		}
	}
}

// End of the discharge phases: the termination test follows.
__global__ void k_ctl_discharge_end(SolveControl *ctl, int phases)
{
	if (threadIdx.x == 0 && blockIdx.x == 0 && ctl->mode == kModeDischarge) {
		ctl->phases += phases;
		ctl->mode = kModeClosureNew;
	}
}

// End of a closure stage. The termination test that reached t goes back to a global relabel; one
// that did not, with no dirty tile left, has found S_min (so has the final closure).
__global__ __launch_bounds__(kCtlThreads) void k_ctl_closure_end(SolveControl *ctl, const uint8_t *dirty, int ntiles)
{
	__shared__ int sums[kCtlThreads / 64];
	const int m = *(volatile const int32_t *)&ctl->mode;
	if (m < kModeClosureNew || m > kModeFinalCont) {
		return;
	}
	const bool check = closure_check_mode(m);
	if (check && *(volatile const int32_t *)&ctl->reached) {
		if (threadIdx.x == 0) {
			ctl->mode = kModeRelabelNew;
		}
		return;
	}
	const int nd = ctl_count_flags(dirty, ntiles, sums);
	if (threadIdx.x == 0) {
		ctl->dirty = nd;
		if (nd > 0) {
			ctl->mode = check ? kModeClosureCont : kModeFinalCont;
		} else {
			ctl->mode = check ? kModeDoneClosed : kModeDoneExact;
		}
	}
}

// Image -> graph front end: t-links by intensity, then marker and border forcing; n-links by the
// absolute intensity difference per neighbour class (the caller-side construction of
// benchmarks/graph_cut/graph_cut.cpp:32-98, as table lookups).
template <int kConn>
__global__ __launch_bounds__(256) void k_caps_from_volume(int nx, int ny, int nz, const uint8_t *__restrict__ vol,
                                                          const uint8_t *__restrict__ markers,
                                                          const cgp_volume_weights *__restrict__ w,
                                                          int32_t *__restrict__ cs, int32_t *__restrict__ ct,
                                                          int32_t *__restrict__ cn)
{
	using D = Dirs<kConn>;
	const int64_t n = (int64_t)nx * ny * nz;
	const int64_t stride = (int64_t)gridDim.x * blockDim.x;
	for (int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; v < n; v += stride) {
		const int x = (int)(v % nx);
		const int y = (int)((v / nx) % ny);
		const int z = (int)(v / ((int64_t)nx * ny));
		const int iv = vol[v];
		int32_t s = w->tlink_source[iv], t = w->tlink_sink[iv];
		if (markers) {
			const int m = markers[v];
			if (m == w->marker_source) {
				s = w->marker_capacity;
				t = 0;
			} else if (m == w->marker_sink) {
				s = 0;
				t = w->marker_capacity;
			}
		}
		const int b = w->border;
		if (b > 0 && (x < b || y < b || z < b || x >= nx - b || y >= ny - b || z >= nz - b)) {
			s = w->marker_capacity;
			t = 0;
		}
		cs[v] = s;
		ct[v] = t;
#pragma unroll
		for (int j = 0; j < D::H; ++j) {
			const int dx = D::dx(j), dy = D::dy(j), dz = D::dz(j);
			const int ux = x + dx, uy = y + dy, uz = z + dz;
			int32_t c = 0;
			if (ux >= 0 && ux < nx && uy >= 0 && uy < ny && uz >= 0 && uz < nz) {
				const int iu = vol[((int64_t)uz * ny + uy) * nx + ux];
				const int cls = (dx != 0) + (dy != 0) + (dz != 0) - 1;
				c = w->nlink[cls][iv > iu ? iv - iu : iu - iv];
			}
			cn[(int64_t)j * n + v] = c;
		}
	}
}

// Saturated-direction bits of every voxel of one grid: bit d when the edge towards direction d
// has capacity > 0 and residual 0 (capacities as in the init kernel).
template <int kConn>
__global__ __launch_bounds__(256) void k_saturated_bits(int nx, int ny, int nz, const int32_t *__restrict__ r,
                                                        const int32_t *__restrict__ cn,
                                                        const int32_t *__restrict__ cnb, uint32_t *__restrict__ out)
{
	using D = Dirs<kConn>;
	const int64_t n = (int64_t)nx * ny * nz;
	const int64_t stride = (int64_t)gridDim.x * blockDim.x;
	for (int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; v < n; v += stride) {
		const int x = (int)(v % nx);
		const int y = (int)((v / nx) % ny);
		const int z = (int)(v / ((int64_t)nx * ny));
		uint32_t bits = 0;
#pragma unroll
		for (int d = 0; d < D::K; ++d) {
			const int ux = x + D::dx(d), uy = y + D::dy(d), uz = z + D::dz(d);
			if (ux < 0 || ux >= nx || uy < 0 || uy >= ny || uz < 0 || uz >= nz) {
				continue;
			}
			int32_t cap;
			if (d < D::H) {
				cap = cn[(int64_t)d * n + v];
			} else {
				const int64_t u = ((int64_t)uz * ny + uy) * nx + ux;
				cap = (cnb ? cnb : cn)[(int64_t)(d - D::H) * n + u];
			}
			if (cap > 0 && r[(int64_t)d * n + v] == 0) {
				bits |= 1u << d;
			}
		}
		out[v] = bits;
	}
}

// Capacity of a cut given by its source-side mask, over one grid (int64 block sums, one atomic per
// block).
template <int kConn>
__global__ __launch_bounds__(256) void k_cut_capacity(int nx, int ny, int nz, const uint8_t *__restrict__ side,
                                                      const int32_t *__restrict__ cs, const int32_t *__restrict__ ct,
                                                      const int32_t *__restrict__ cn,
                                                      const int32_t *__restrict__ cnb,
                                                      unsigned long long *__restrict__ out)
{
	using D = Dirs<kConn>;
	__shared__ long long red[4];
	const int64_t n = (int64_t)nx * ny * nz;
	const int64_t stride = (int64_t)gridDim.x * blockDim.x;
	long long sum = 0;
	for (int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; v < n; v += stride) {
		const int x = (int)(v % nx);
		const int y = (int)((v / nx) % ny);
		const int z = (int)(v / ((int64_t)nx * ny));
		const bool in = side[v] != 0;
		sum += in ? ct[v] : cs[v];
#pragma unroll
		for (int j = 0; j < D::H; ++j) {
			const int ux = x + D::dx(j), uy = y + D::dy(j), uz = z + D::dz(j);
			if (ux < 0 || ux >= nx || uy < 0 || uy >= ny || uz < 0 || uz >= nz) {
				continue;
			}
			const bool uin = side[((int64_t)uz * ny + uy) * nx + ux] != 0;
			if (in && !uin) {
				sum += cn[(int64_t)j * n + v];            // v -> v + o_j leaves the source side
			} else if (uin && !in) {
				sum += (cnb ? cnb : cn)[(int64_t)j * n + v];  // v + o_j -> v
			}
		}
	}
	const long long total = block_sum_i64<256>(sum, red);
	if (threadIdx.x == 0 && total) {
		atomicAdd(out + (blockIdx.x & (kSlots - 1)) * kSlotStride, (unsigned long long)total);  // slotted: see kSlots
	}
}

__global__ __launch_bounds__(256) void k_fill_u8(uint8_t *__restrict__ p, int64_t n, uint8_t value,
                                                 const SolveControl *ctl = nullptr, unsigned gate = 0)
{
	if (!gate_open(ctl, gate)) {
		return;
	}
	const int64_t stride = (int64_t)gridDim.x * blockDim.x;
	for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
		p[i] = value;
	}
}

__global__ __launch_bounds__(256) void k_fill_i32(int32_t *__restrict__ p, int64_t n, int32_t value,
                                                  const SolveControl *ctl = nullptr, unsigned gate = 0)
{
	if (!gate_open(ctl, gate)) {
		return;
	}
	const int64_t stride = (int64_t)gridDim.x * blockDim.x;
	for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
		p[i] = value;
	}
}

// Start of a solve, one launch: the init status, the slotted visit counters and flows (one
// region), the alternating counts of the multi-workgroup compactions and, for the device-driven
// solve, the control block (cursor_words must not exceed the threads of the launch).
__global__ __launch_bounds__(256) void k_solve_reset(int32_t *__restrict__ status, unsigned long long *__restrict__ slots,
                                                     int64_t slot_words, int32_t *__restrict__ cursors, int cursor_words,
                                                     SolveControl *ctl, int32_t max_rounds)
{
	const int64_t i0 = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
	const int64_t stride = (int64_t)gridDim.x * blockDim.x;
	for (int64_t i = i0; i < slot_words; i += stride) {
		slots[i] = 0;
	}
	if (i0 < cursor_words) {
		cursors[i0] = 0;
	}
	if (i0 == 0) {
		*status = 0;
		if (ctl) {
			SolveControl c = {};
			c.mode = kModeInit;
			c.max_rounds = max_rounds;
			*ctl = c;
		}
	}
}

// Start of an S_min closure, one launch: s[0, n) cleared with 16-byte stores (s is 16-byte
// aligned), and, when given, the reached flag and the closure's tile flags dirty[0, ntiles).
__global__ __launch_bounds__(256) void k_closure_reset(uint8_t *__restrict__ s, int64_t n, int32_t *__restrict__ reached,
                                                       uint8_t *__restrict__ dirty, int64_t ntiles,
                                                       const SolveControl *ctl, unsigned gate)
{
	if (!gate_open(ctl, gate)) {
		return;
	}
	const int64_t i0 = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
	const int64_t stride = (int64_t)gridDim.x * blockDim.x;
	const int64_t words = n / 16;
	uint4 *s4 = reinterpret_cast<uint4 *>(s);
	for (int64_t i = i0; i < words; i += stride) {
		s4[i] = make_uint4(0u, 0u, 0u, 0u);
	}
	for (int64_t i = words * 16 + i0; i < n; i += stride) {
		s[i] = 0;
	}
	if (dirty) {
		for (int64_t i = i0; i < ntiles; i += stride) {
			dirty[i] = 0;
		}
	}
	if (reached && i0 == 0) {
		*reached = 0;
	}
}

// Set flags of a[0, n) and b[0, n) (b may be null) in one launch over many workgroups (4096 flags
// each, 4 a thread by one 4-byte load; the arrays are padded to whole 32-byte pieces): every
// workgroup adds its counts into acc[0], acc[1] and takes a ticket in acc[2]; the last one stores
// the totals in out[0], out[1] and clears acc for the next launch (acc starts at zero).
__global__ __launch_bounds__(1024) void k_count_flags2(const uint8_t *__restrict__ a, const uint8_t *__restrict__ b,
                                                       int n, int32_t *__restrict__ out, int32_t *__restrict__ acc,
                                                       const unsigned long long *__restrict__ slots, int64_t nsum,
                                                       unsigned long long *__restrict__ total)
{
	__shared__ int sums[2][1024 / 64];
	__shared__ int last;
	const int i = ((int)blockIdx.x * 1024 + (int)threadIdx.x) * 4;
	int ca = 0, cb = 0;
	if (i < n) {
		const uint32_t wa = *reinterpret_cast<const uint32_t *>(a + i);
		const uint32_t wb = b ? *reinterpret_cast<const uint32_t *>(b + i) : 0u;
#pragma unroll
		for (int j = 0; j < 4; ++j) {
			const bool in = i + j < n;
			ca += (in && ((wa >> (8 * j)) & 0xffu)) ? 1 : 0;
			cb += (in && ((wb >> (8 * j)) & 0xffu)) ? 1 : 0;
		}
	}
#pragma unroll
	for (int off = 32; off > 0; off >>= 1) {
		ca += __shfl_xor(ca, off, 64);
		cb += __shfl_xor(cb, off, 64);
	}
	const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
	if (lane == 0) {
		sums[0][wave] = ca;
		sums[1][wave] = cb;
	}
	__syncthreads();
	if (threadIdx.x == 0) {
		int ta = 0, tb = 0;
		for (int w = 0; w < 1024 / 64; ++w) {
			ta += sums[0][w];
			tb += sums[1][w];
		}
		if (ta) {
			atomicAdd(&acc[0], ta);
		}
		if (tb) {
			atomicAdd(&acc[1], tb);
		}
		__threadfence();
		last = 0;
		if (atomicAdd(&acc[2], 1) == (int)gridDim.x - 1) {
			__threadfence();
			out[0] = atomicExch(&acc[0], 0);
			out[1] = atomicExch(&acc[1], 0);
			atomicExch(&acc[2], 0);
			last = 1;
		}
	}
	if (slots) {
		// the last block also sums the slotted visit counters and flows (k_sum_slots' work; earlier
		// launches wrote them)
		__syncthreads();
		if (last) {
			for (int64_t i = threadIdx.x; i < nsum; i += blockDim.x) {
				unsigned long long sum = 0;
#pragma unroll
				for (int j = 0; j < kSlots; ++j) {
					sum += slots[i * kSlotWords + j * kSlotStride];
				}
				total[i] = sum;
			}
		}
	}
}

__global__ __launch_bounds__(256) void k_sink_side_mask(const int32_t *__restrict__ h, int64_t n,
                                                        uint8_t *__restrict__ out, uint8_t mask, uint8_t background)
{
	const int64_t stride = (int64_t)gridDim.x * blockDim.x;
	for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
		out[i] = h[i] < kLabelInf ? mask : background;
	}
}

}  // namespace cgp

#include "grid_wave_tiles.h"
