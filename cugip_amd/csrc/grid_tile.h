// Compile-time geometry of the 3D grid path: neighbourhood directions and the
// LDS tile (own voxels plus a one-voxel halo) that every grid kernel stages.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <type_traits>

namespace cgp {

// Directed residual directions. d < H: the negative half neighbourhood in the
// reference's table order (VonNeumannNeighborhood<3>::offset(1..3),
// MooreNeighborhood<3>::offset(1..13), cugip/neighborhood.hpp); d >= H: mirror.
template <int kConn>
struct Dirs;

template <>
struct Dirs<6> {
	static constexpr int K = 6;
	static constexpr int H = 3;
	__host__ __device__ static constexpr int dx(int d) { return d == 0 ? -1 : (d == 3 ? 1 : 0); }
	__host__ __device__ static constexpr int dy(int d) { return d == 1 ? -1 : (d == 4 ? 1 : 0); }
	__host__ __device__ static constexpr int dz(int d) { return d == 2 ? -1 : (d == 5 ? 1 : 0); }
};

template <>
struct Dirs<26> {
	static constexpr int K = 26;
	static constexpr int H = 13;
	__host__ __device__ static constexpr int hx(int h) { return h % 3 - 1; }
	__host__ __device__ static constexpr int hy(int h) { return (h / 3) % 3 - 1; }
	__host__ __device__ static constexpr int hz(int h) { return h / 9 - 1; }
	__host__ __device__ static constexpr int dx(int d) { return d < H ? hx(d) : -hx(d - H); }
	__host__ __device__ static constexpr int dy(int d) { return d < H ? hy(d) : -hy(d - H); }
	__host__ __device__ static constexpr int dz(int d) { return d < H ? hz(d) : -hz(d - H); }
};

template <int kConn>
__host__ __device__ constexpr int opposite(int d)
{
	return d < Dirs<kConn>::H ? d + Dirs<kConn>::H : d - Dirs<kConn>::H;
}

// Tile of TX x TY x TZ voxels processed by one workgroup of TX*TY threads; thread
// (lx, ly) owns the z-column lz = 0..TZ-1. Labels are staged with a halo of one
// voxel on every side (PX x PY x PZ).
template <int kConn, int kTX, int kTY, int kTZ>
struct Tile {
	using D = Dirs<kConn>;
	static constexpr int K = D::K;
	static constexpr int TX = kTX, TY = kTY, TZ = kTZ;
	static constexpr int NT = TX * TY;
	static constexpr int NV = TX * TY * TZ;
	static constexpr int PX = TX + 2, PY = TY + 2, PZ = TZ + 2;
	static constexpr int NP = PX * PY * PZ;
	static_assert(NT % 64 == 0, "a tile must map to whole wave64s");

	__host__ __device__ static constexpr int poff(int d) { return (D::dz(d) * PY + D::dy(d)) * PX + D::dx(d); }
	__host__ __device__ static constexpr int voff(int d) { return (D::dz(d) * TY + D::dy(d)) * TX + D::dx(d); }
	__host__ __device__ static constexpr int pidx(int x, int y, int z) { return ((z + 1) * PY + (y + 1)) * PX + (x + 1); }
	__host__ __device__ static constexpr int vidx(int x, int y, int z) { return (z * TY + y) * TX + x; }
};

// One grid of a batch. Per-voxel arrays of all instances are packed one after another: the
// instance's voxels start at vbase, its residual planes at conn * vbase (plane stride nvox) and
// its n-link planes at (conn / 2) * vbase. Tiles of all instances share one id space.
struct InstanceDesc {
	int32_t nx, ny, nz;
	int32_t ntx, nty, ntz;
	int32_t tbase;  // global id of the instance's first tile
	int32_t conn;
	int64_t vbase;
	int64_t nvox;
};

// Everything a grid kernel needs, passed by value.
struct GridView {
	int32_t nx, ny, nz;
	int32_t ntx, nty, ntz;   // tiles per axis
	int64_t nvox;            // voxels = stride between residual planes
	int32_t *e;              // excess
	int32_t *rt;             // residual sink t-link
	int32_t *h;              // label: BFS distance to t, kLabelInf if t is unreachable
	int32_t *r;              // K residual planes, r[d * nvox + v] = residual v -> v + o_d
	uint8_t *s;              // S_min byte mask
	uint32_t *omask;         // state word: open directions (r > 0), residual sink link > 0, excess > 0
	const int32_t *cs, *ct;  // input t-link capacities
	const int32_t *cn, *cnb; // input n-link capacities, H planes (cnb may be null)
	unsigned long long *flow;  // kSlots slots of this instance's flow into t
	int32_t *status;         // nonzero: input error detected on the device
	unsigned long long *counters; // visit counters, kSlots slots per kind (see counter_slot)
	const InstanceDesc *inst;  // batch: instance table (null for a single grid)
	const int32_t *tile_inst;  // batch: instance of every global tile
	int32_t tbase;             // global id of this view's first tile
};

// Work classes of the ordered discharge lists (heaviest first).
constexpr int kWorkClasses = 8;

// The flow of an instance and the visit counters are sums over kSlots slots, each on its own
// 128-byte line, picked by tile id: thousands of workgroups adding into one address serialise in
// the L2 and hold up every other request queued behind them.
constexpr int kSlots = 32;
constexpr int kSlotStride = 16;  // unsigned long longs per slot (128 bytes)
constexpr int kSlotWords = kSlots * kSlotStride;  // per instance (flow) or per counter kind

__device__ inline unsigned long long *flow_slot(const GridView &g, int tile)
{
	return g.flow + (tile & (kSlots - 1)) * kSlotStride;
}

// kind: 0 discharge visits, 1 relabel visits, 2 closure visits
__device__ inline unsigned long long *counter_slot(const GridView &g, int kind, int tile)
{
	return g.counters + kind * kSlotWords + (tile & (kSlots - 1)) * kSlotStride;
}

// The view of instance i of a batch.
__device__ inline GridView instance_at(const GridView &g, int i)
{
	const InstanceDesc d = g.inst[i];
	GridView v = g;
	v.nx = d.nx;
	v.ny = d.ny;
	v.nz = d.nz;
	v.ntx = d.ntx;
	v.nty = d.nty;
	v.ntz = d.ntz;
	v.nvox = d.nvox;
	v.tbase = d.tbase;
	v.e = g.e + d.vbase;
	v.rt = g.rt + d.vbase;
	v.h = g.h + d.vbase;
	v.s = g.s + d.vbase;
	v.omask = g.omask + d.vbase;
	v.cs = g.cs + d.vbase;
	v.ct = g.ct + d.vbase;
	v.r = g.r + d.vbase * d.conn;
	v.cn = g.cn + d.vbase * (d.conn / 2);
	v.cnb = g.cnb ? g.cnb + d.vbase * (d.conn / 2) : nullptr;
	v.flow = g.flow + (int64_t)i * kSlotWords;
	return v;
}

// The view of the instance that owns global tile `tile`, and the tile's id inside it. For a single
// grid this is the view itself.
__device__ inline GridView instance_view(const GridView &g, int tile, int &local)
{
	if (!g.inst) {
		local = tile;
		return g;
	}
	const GridView v = instance_at(g, g.tile_inst[tile]);
	local = tile - v.tbase;
	return v;
}

__device__ inline int64_t voxel_index(const GridView &g, int x, int y, int z)
{
	return ((int64_t)z * g.ny + y) * g.nx + x;
}

// Element offset of z-slice z in a per-voxel array. Within a workgroup z is uniform wherever this is
// used, so the slice start lives in scalar registers.
__device__ inline int64_t slice_start(const GridView &g, int z)
{
	return (int64_t)z * ((int64_t)g.nx * g.ny);
}

// In-slice index of (x, y).
__device__ inline uint32_t slice_xy(const GridView &g, int x, int y)
{
	return (uint32_t)(y * g.nx + x);
}

// Element `off` after `base`, addressed as a scalar base + a 32-bit byte offset per lane instead of
// 64-bit index arithmetic per lane. Callers keep off * sizeof(P) below 2^32: the grid setup limits
// a z-slice to fewer than 2^26 voxels and the offsets below span at most TZ + 1 slices.
template <typename P>
__device__ inline P &at(P *base, uint32_t off)
{
	using B = typename std::conditional<std::is_const<P>::value, const char, char>::type;
	return *reinterpret_cast<P *>(reinterpret_cast<B *>(base) + off * (uint32_t)sizeof(P));
}

// Voxel xy of z-slice oz + k of a per-voxel array (k a compile-time constant in [-1, TZ] after
// unrolling): for k >= 0 the tile's first slice is the base, the slice below the tile has its own.
template <typename P>
__device__ inline P &tile_at(P *array, const GridView &g, int oz, int k, uint32_t xy, uint32_t sxy)
{
	if (k < 0) {
		return at(array + slice_start(g, oz + k), xy);
	}
	return at(array + slice_start(g, oz), xy + (uint32_t)k * sxy);
}

__device__ inline bool inside(const GridView &g, int x, int y, int z)
{
	return x >= 0 && x < g.nx && y >= 0 && y < g.ny && z >= 0 && z < g.nz;
}

__device__ inline void tile_coords(const GridView &g, int tile, int &tx, int &ty, int &tz)
{
	tx = tile % g.ntx;
	const int rest = tile / g.ntx;
	ty = rest % g.nty;
	tz = rest / g.nty;
}

}  // namespace cgp
