// Wave-per-tile kernels of the 6-connected grid path (gfx950, wave64).
//
// One workgroup = one wave of 64 lanes = one tile of 32 x 8 x 4 voxels. The tile's voxels form
// R = 32 rows (row r = z*8 + y) of 32 voxels; a row is one 32-bit word of every bit plane:
//   plane d = 0..5  residual in direction d is > 0 ("open"; d as in Dirs<6>)
//   plane 6         residual sink t-link > 0 ("seed": level 1 of the BFS)
//   plane 7         excess > 0
// Bit planes live in HBM tile-major (plane, tile, row), so a tile reads each plane as 128
// contiguous bytes. Lanes 0..31 ("row lanes") run the BFS on whole rows with shifts and
// cross-lane reads; all 64 lanes ("voxel lanes": lane l owns x = l & 31 of rows 2k + (l >> 5),
// k = 0..15) move the int32 per-voxel state between HBM and LDS with coalesced accesses.
//
// Labels. A tile only ever reads the labels of the voxels on its neighbours' faces, so the
// global relabel stores labels for the 832 face voxels of each tile (tile-major, contiguous per
// face) plus one "reaches t" bit plane; interior labels exist only inside a kernel.
//
// Tile BFS with exits: a voxel's label is the smallest of 1 (seed), 1 + label of an open halo
// neighbour (an "exit", only on faces) and 1 + label of an open in-tile neighbour. It runs level by
// level from the lowest source: F_L = (expand(F_{L-1}) | sources_at(L)) & ~visited, with the
// sources binned into a window of kWindow levels (re-binned when the BFS runs past the window or
// jumps over empty levels). The same routine serves the global relabel (halo = current global
// face labels) and the region relabel of a discharge visit.
#pragma once

#include "common.h"
#include "grid_tile.h"

namespace cgp {
namespace wave {

constexpr int TX = 32, TY = 8, TZ = 4;
constexpr int R = TY * TZ;   // 32 rows
constexpr int NV = TX * R;   // 1024 voxels
constexpr int K = 6;
constexpr int kPlanes = 8;
constexpr int kSeedPlane = 6, kExcessPlane = 7;
constexpr int kWindow = 32;  // source levels binned per window
constexpr int kStore = 32;   // levels swept per push chunk (one more level is kept below it)
constexpr int kVoxPerLane = NV / 64;  // 16

// Face entries of a tile, in this order: x = 0 (per row), x = 31 (per row), y = 0 (per z, x),
// y = 7 (per z, x), z = 0 (per y, x), z = 3 (per y, x).
constexpr int kFXM = 0, kFXP = 32, kFYM = 64, kFYP = 192, kFZM = 320, kFZP = 576, kFaces = 832;

__device__ inline int row_y(int r) { return r % TY; }
__device__ inline int row_z(int r) { return r / TY; }
// voxel lane mapping: lane l, slot k -> row 2k + (l >> 5), x = l & 31
__device__ inline int slot_row(int lane, int k) { return 2 * k + (lane >> 5); }

__device__ inline size_t plane_word(int plane, int tile, int row, int ntiles)
{
	return ((size_t)plane * ntiles + tile) * R + row;
}

// Direction (Dirs<6>) that leaves the tile through face entry t.
__device__ inline int face_dir(int t)
{
	return t < kFXP ? 0 : t < kFYM ? 3 : t < kFYP ? 1 : t < kFZM ? 4 : t < kFZP ? 2 : 5;
}

// Voxel (row r, bit x) of face entry t.
__device__ inline void face_voxel(int t, int &r, int &x)
{
	if (t < kFXP) {
		r = t;
		x = 0;
	} else if (t < kFYM) {
		r = t - kFXP;
		x = TX - 1;
	} else if (t < kFYP) {
		const int i = t - kFYM;
		r = (i >> 5) * TY;
		x = i & 31;
	} else if (t < kFZM) {
		const int i = t - kFYP;
		r = (i >> 5) * TY + TY - 1;
		x = i & 31;
	} else if (t < kFZP) {
		const int i = t - kFZM;
		r = i >> 5;
		x = i & 31;
	} else {
		const int i = t - kFZP;
		r = (TZ - 1) * TY + (i >> 5);
		x = i & 31;
	}
}

// Face entry of the voxel (row r, bit x) for direction d leaving the tile.
__device__ inline int face_entry(int d, int r, int x)
{
	switch (d) {
	case 0: return kFXM + r;
	case 3: return kFXP + r;
	case 1: return kFYM + row_z(r) * TX + x;
	case 4: return kFYP + row_z(r) * TX + x;
	case 2: return kFZM + row_y(r) * TX + x;
	default: return kFZP + row_y(r) * TX + x;
	}
}

// Does direction d leave the tile from in-tile voxel (row r, bit x)?
__device__ inline bool leaves(int x, int r, int d)
{
	switch (d) {
	case 0: return x == 0;
	case 3: return x == TX - 1;
	case 1: return row_y(r) == 0;
	case 4: return row_y(r) == TY - 1;
	case 2: return row_z(r) == 0;
	default: return row_z(r) == TZ - 1;
	}
}

// In-tile neighbour index (valid only when !leaves()).
__device__ inline int neighbour_index(int i, int d)
{
	switch (d) {
	case 0: return i - 1;
	case 3: return i + 1;
	case 1: return i - TX;
	case 4: return i + TX;
	case 2: return i - TX * TY;
	default: return i + TX * TY;
	}
}

// Neighbour tile slot (0..26) for direction d of Dirs<6>.
__device__ inline int dir_slot(int d)
{
	using D = Dirs<6>;
	return ((D::dz(d) + 1) * 3 + (D::dy(d) + 1)) * 3 + (D::dx(d) + 1);
}

__device__ inline void tile_origin(const GridView &g, int tile, int &tx, int &ty, int &tz, int &ox, int &oy, int &oz)
{
	tile_coords(g, tile, tx, ty, tz);
	ox = tx * TX;
	oy = ty * TY;
	oz = tz * TZ;
}

// Halo labels: halo[t] = global label of the neighbour across face entry t, read from the
// opposite face of the neighbouring tile (kLabelInf outside the tile grid).
__device__ inline void load_halo(const GridView &g, int32_t *halo, int tx, int ty, int tz, int lane)
{
	for (int t = lane; t < kFaces; t += 64) {
		int ntx = tx, nty = ty, ntz = tz, src;
		if (t < kFXP) {
			ntx = tx - 1;
			src = kFXP + t;
		} else if (t < kFYM) {
			ntx = tx + 1;
			src = t - kFXP;
		} else if (t < kFYP) {
			nty = ty - 1;
			src = kFYP + (t - kFYM);
		} else if (t < kFZM) {
			nty = ty + 1;
			src = kFYM + (t - kFYP);
		} else if (t < kFZP) {
			ntz = tz - 1;
			src = kFZP + (t - kFZM);
		} else {
			ntz = tz + 1;
			src = kFZM + (t - kFZP);
		}
		int32_t v = kLabelInf;
		if (ntx >= 0 && ntx < g.ntx && nty >= 0 && nty < g.nty && ntz >= 0 && ntz < g.ntz) {
			v = g.faces[(size_t)((ntz * g.nty + nty) * g.ntx + ntx) * kFaces + src];
		}
		halo[t] = v;
	}
}

// Cross-lane read of a row word.
__device__ inline uint32_t lane_read(uint32_t v, int src)
{
	return (uint32_t)__shfl((int)v, src, 64);
}

// Pull expansion of frontier F for row lane r with open words o[6]: v joins if open_d(v) and its
// neighbour in direction d is in F. Lanes >= R pass zero words.
__device__ inline uint32_t expand_rows(uint32_t F, const uint32_t (&o)[K], int r)
{
	const int y = row_y(r), z = row_z(r);
	const uint32_t fym = lane_read(F, r - 1), fyp = lane_read(F, r + 1);
	const uint32_t fzm = lane_read(F, r - TY), fzp = lane_read(F, r + TY);
	uint32_t n = (o[0] & (F << 1)) | (o[3] & (F >> 1));
	n |= y > 0 ? (o[1] & fym) : 0u;
	n |= y < TY - 1 ? (o[4] & fyp) : 0u;
	n |= z > 0 ? (o[2] & fzm) : 0u;
	n |= z < TZ - 1 ? (o[5] & fzp) : 0u;
	return n;
}

__device__ inline int wave_min_i32(int v)
{
#pragma unroll
	for (int off = 32; off > 0; off >>= 1) {
		const int o = __shfl_xor(v, off, 64);
		v = o < v ? o : v;
	}
	return v;
}

__device__ inline int wave_max_i32(int v)
{
#pragma unroll
	for (int off = 32; off > 0; off >>= 1) {
		const int o = __shfl_xor(v, off, 64);
		v = o > v ? o : v;
	}
	return v;
}

// Exclusive prefix over the 64 lanes; total in `total`.
__device__ inline int wave_exclusive_prefix(int v, int lane, int &total)
{
	int incl = v;
#pragma unroll
	for (int off = 1; off < 64; off <<= 1) {
		const int t = __shfl_up(incl, off, 64);
		if (lane >= off) {
			incl += t;
		}
	}
	total = __shfl(incl, 63, 64);
	return incl - v;
}

// Shared memory of the tile BFS.
struct BfsShared {
	int32_t fexit[kFaces];      // exit level of each face entry (1 + open halo label), kLabelInf if none
	uint32_t inj[kWindow][R];   // sources binned by level - window base
	uint32_t visw[R];           // visited rows (for re-binning)
};

// Level sets kept by a discharge for its push sweep: level indices [lo, lo + kStore + 1).
struct LevelStore {
	uint32_t ls[kStore + 1][R];
	int32_t lvl[kStore + 1];
};

// Exit level of every face entry from the halo labels and the tile's open words (LDS).
__device__ inline void compute_exits(BfsShared &bs, const int32_t *halo, const uint32_t (*openw)[R], int lane)
{
	for (int t = lane; t < kFaces; t += 64) {
		int r, x;
		face_voxel(t, r, x);
		const int d = face_dir(t);
		int32_t e = kLabelInf;
		if ((openw[d][r] >> x) & 1u) {
			const int32_t h = halo[t];
			e = h < kLabelInf ? h + 1 : kLabelInf;
		}
		bs.fexit[t] = e;
	}
}

// Bin the sources (seeds at level 1, face exits) of unvisited voxels with level in
// [base, base + kWindow); returns the highest binned level and the lowest level beyond the window.
__device__ inline void bin_sources(BfsShared &bs, int base, uint32_t seed, uint32_t valid, uint32_t vis, int lane,
                                   int &last_binned, int &next_pending)
{
	for (int t = lane; t < kWindow * R; t += 64) {
		(&bs.inj[0][0])[t] = 0u;
	}
	__syncthreads();
	int hi = -1, pend = kLabelInf;
	if (lane < R && base <= 1 && 1 < base + kWindow) {
		const uint32_t s = seed & valid & ~vis;
		if (s) {
			atomicOr(&bs.inj[1 - base][lane], s);
			hi = 1;
		}
	}
	for (int t = lane; t < kFaces; t += 64) {
		const int32_t e = bs.fexit[t];
		if (e >= kLabelInf || e < base) {
			continue;
		}
		int r, x;
		face_voxel(t, r, x);
		if ((bs.visw[r] >> x) & 1u) {
			continue;
		}
		if (e < base + kWindow) {
			atomicOr(&bs.inj[e - base][r], 1u << x);
			hi = e > hi ? e : hi;
		} else {
			pend = e < pend ? e : pend;
		}
	}
	__syncthreads();
	last_binned = wave_max_i32(hi);
	next_pending = wave_min_i32(pend);
}

// Record face labels of the newly visited voxels Fn of row r at `level`.
__device__ inline void record_faces(int32_t *faceout, uint32_t Fn, int r, int level)
{
	const int y = row_y(r), z = row_z(r);
	if (Fn & 1u) {
		faceout[kFXM + r] = level;
	}
	if (Fn >> 31) {
		faceout[kFXP + r] = level;
	}
	const bool yface = y == 0 || y == TY - 1, zface = z == 0 || z == TZ - 1;
	if (!yface && !zface) {
		return;
	}
	uint32_t bits = Fn;
	while (bits) {
		const int b = __ffs((int)bits) - 1;
		bits &= bits - 1u;
		if (y == 0) {
			faceout[kFYM + z * TX + b] = level;
		} else if (y == TY - 1) {
			faceout[kFYP + z * TX + b] = level;
		}
		if (z == 0) {
			faceout[kFZM + y * TX + b] = level;
		} else if (z == TZ - 1) {
			faceout[kFZP + y * TX + b] = level;
		}
	}
}

// Tile BFS. Row lanes (lane < R) supply their open, seed and valid words (lanes >= R pass zero).
// Returns the number of non-empty levels and (row lanes) the visited words. Optional outputs:
// faceout[kFaces] (labels of face voxels), lab[NV] (labels of all voxels), store (level sets with
// level index in [store_lo, store_lo + kStore]). Unreached voxels keep kLabelInf.
__device__ inline int tile_bfs(BfsShared &bs, LevelStore *store, const uint32_t (&o)[K], uint32_t seed, uint32_t valid,
                               int32_t *faceout, int32_t *lab, int store_lo, int lane, uint32_t &visited_out)
{
	const bool row_lane = lane < R;
	int lo = __ballot(row_lane && (seed & valid) != 0u) != 0ull ? 1 : kLabelInf;
	{
		int m = kLabelInf;
		for (int t = lane; t < kFaces; t += 64) {
			const int32_t e = bs.fexit[t];
			m = e < m ? e : m;
		}
		m = wave_min_i32(m);
		lo = m < lo ? m : lo;
	}
	if (faceout) {
		for (int t = lane; t < kFaces; t += 64) {
			faceout[t] = kLabelInf;
		}
	}
	if (lab) {
		for (int t = lane; t < NV; t += 64) {
			lab[t] = kLabelInf;
		}
	}
	if (row_lane) {
		bs.visw[lane] = 0u;
	}
	uint32_t vis = 0u, F = 0u;
	int count = 0;
	int base = lo;
	while (base < kLabelInf) {
		int last_binned, next_pending;
		__syncthreads();
		bin_sources(bs, base, seed, valid, vis, lane, last_binned, next_pending);
		bool window_done = true;
		for (int w = 0; w < kWindow; ++w) {
			const int level = base + w;
			const uint32_t injw = row_lane ? bs.inj[w][lane] : 0u;
			const uint32_t Fn = (expand_rows(F, o, lane) | injw) & valid & ~vis;
			F = Fn;
			if (__ballot(Fn != 0u) == 0ull) {
				if (level >= last_binned) {
					window_done = false;  // frontier empty and nothing left in this window
					break;
				}
				continue;
			}
			vis |= Fn;
			if (store && count >= store_lo && count <= store_lo + kStore) {
				if (row_lane) {
					store->ls[count - store_lo][lane] = Fn;
				}
				if (lane == 0) {
					store->lvl[count - store_lo] = level;
				}
			}
			if (row_lane && faceout) {
				record_faces(faceout, Fn, lane, level);
			}
			if (row_lane && lab) {
				uint32_t bits = Fn;
				while (bits) {
					const int b = __ffs((int)bits) - 1;
					lab[lane * TX + b] = level;
					bits &= bits - 1u;
				}
			}
			++count;
		}
		if (row_lane) {
			bs.visw[lane] = vis;
		}
		const bool growing = window_done && __ballot(F != 0u) != 0ull;
		base = growing ? base + kWindow : next_pending;
	}
	__syncthreads();
	visited_out = vis;
	return count;
}

// Row word of valid voxels (inside the volume) for row r of the tile at (ox, oy, oz).
__device__ inline uint32_t valid_word(const GridView &g, int ox, int oy, int oz, int r)
{
	if (r >= R || oy + row_y(r) >= g.ny || oz + row_z(r) >= g.nz) {
		return 0u;
	}
	const int w = g.nx - ox;
	return w >= TX ? 0xffffffffu : ((1u << w) - 1u);
}

// Assemble row words from a per-voxel predicate of slot k: rows 2k (lanes 0..31) and 2k+1
// (lanes 32..63) of the ballot land in row lanes 2k and 2k+1.
__device__ inline void ballot_rows(bool pred, int k, int lane, uint32_t &word)
{
	const unsigned long long b = __ballot(pred);
	if (lane == 2 * k) {
		word = (uint32_t)b;
	} else if (lane == 2 * k + 1) {
		word = (uint32_t)(b >> 32);
	}
}

// ---------------------------------------------------------------------------------------------
// init: residual planes, cancellation min(cs, ct) into the flow, excess, residual sink link and
// the tile's bit planes. Status bit 0: negative capacity; bit 1: int32 headroom exceeded.
__global__ __launch_bounds__(64) void k_wave_init(GridView g, uint32_t *__restrict__ bits, int ntiles)
{
	using D = Dirs<6>;
	const int tile = blockIdx.x, lane = threadIdx.x;
	int tx, ty, tz, ox, oy, oz;
	tile_origin(g, tile, tx, ty, tz, ox, oy, oz);
	const int x = lane & 31;
	long long cancelled = 0;
	bool negative = false, overflow = false;
	uint32_t words[kPlanes] = {0, 0, 0, 0, 0, 0, 0, 0};
	for (int k = 0; k < kVoxPerLane; ++k) {
		const int r = slot_row(lane, k);
		const int gx = ox + x, gy = oy + row_y(r), gz = oz + row_z(r);
		const bool valid = inside(g, gx, gy, gz);
		int32_t res[K] = {0, 0, 0, 0, 0, 0};
		int32_t ev = 0, rtv = 0;
		if (valid) {
			const int64_t v = voxel_index(g, gx, gy, gz);
			const int32_t cs = g.cs[v], ct = g.ct[v];
			negative |= (cs < 0) | (ct < 0);
			const int32_t m = cs < ct ? cs : ct;
			cancelled += m;
			long long incoming = cs - m;
#pragma unroll
			for (int d = 0; d < K; ++d) {
				const int ux = gx + D::dx(d), uy = gy + D::dy(d), uz = gz + D::dz(d);
				int32_t out = 0, in = 0;
				if (inside(g, ux, uy, uz)) {
					if (d < D::H) {
						out = g.cn[(int64_t)d * g.nvox + v];
						in = g.cnb ? g.cnb[(int64_t)d * g.nvox + v] : out;
					} else {
						const int64_t u = voxel_index(g, ux, uy, uz);
						in = g.cn[(int64_t)(d - D::H) * g.nvox + u];
						out = g.cnb ? g.cnb[(int64_t)(d - D::H) * g.nvox + u] : in;
					}
					negative |= (out < 0) | (in < 0);
					overflow |= ((long long)out + (long long)in) > 0x7fffffffLL;
					incoming += in;
				}
				res[d] = out;
				g.r[(int64_t)d * g.nvox + v] = out;
			}
			overflow |= incoming > 0x7fffffffLL;
			ev = cs - m;
			rtv = ct - m;
			g.e[v] = ev;
			g.rt[v] = rtv;
		}
#pragma unroll
		for (int d = 0; d < K; ++d) {
			ballot_rows(res[d] > 0, k, lane, words[d]);
		}
		ballot_rows(rtv > 0, k, lane, words[kSeedPlane]);
		ballot_rows(ev > 0, k, lane, words[kExcessPlane]);
	}
	if (lane < R) {
#pragma unroll
		for (int p = 0; p < kPlanes; ++p) {
			bits[plane_word(p, tile, lane, ntiles)] = words[p];
		}
	}
	if (negative || overflow) {
		atomicOr(g.status, (negative ? 1 : 0) | (overflow ? 2 : 0));
	}
	const long long total = wave_sum_i64(cancelled);
	if (lane == 0 && total) {
		atomicAdd(flow_slot(g, tile), (unsigned long long)total);
	}
}

// Open, seed and excess words of this row lane: from the bit planes, or, with threshold > 0
// (edges with residual <= threshold closed), from the residual planes.
__device__ inline void load_rows(const GridView &g, const uint32_t *bits, int ntiles, int tile, int ox, int oy, int oz,
                                 int lane, int32_t threshold, uint32_t (&o)[K], uint32_t &seed, uint32_t &excess)
{
#pragma unroll
	for (int d = 0; d < K; ++d) {
		o[d] = 0u;
	}
	seed = excess = 0u;
	if (threshold <= 0) {
		if (lane < R) {
#pragma unroll
			for (int d = 0; d < K; ++d) {
				o[d] = bits[plane_word(d, tile, lane, ntiles)];
			}
			seed = bits[plane_word(kSeedPlane, tile, lane, ntiles)];
			excess = bits[plane_word(kExcessPlane, tile, lane, ntiles)];
		}
		return;
	}
	const int x = lane & 31;
	for (int k = 0; k < kVoxPerLane; ++k) {
		const int r = slot_row(lane, k);
		const int gx = ox + x, gy = oy + row_y(r), gz = oz + row_z(r);
		const bool valid = inside(g, gx, gy, gz);
		const int64_t v = valid ? voxel_index(g, gx, gy, gz) : 0;
#pragma unroll
		for (int d = 0; d < K; ++d) {
			ballot_rows(valid && g.r[(int64_t)d * g.nvox + v] > threshold, k, lane, o[d]);
		}
		ballot_rows(valid && g.rt[v] > threshold, k, lane, seed);
		ballot_rows(valid && g.e[v] > 0, k, lane, excess);
	}
	if (lane >= R) {
#pragma unroll
		for (int d = 0; d < K; ++d) {
			o[d] = 0u;
		}
		seed = excess = 0u;
	}
}

// ---------------------------------------------------------------------------------------------
// Global relabel pass (labels only fall). Writes lowered face labels and the reach plane, marks
// the neighbouring tiles whose voxels could now improve, refreshes the tile's active flag. With
// labels_out non-null it also writes every voxel's label (used for the labels accessor only).
__global__ __launch_bounds__(64) void k_wave_bfs(GridView g, const uint32_t *__restrict__ bits, int ntiles,
                                                 const int32_t *__restrict__ list, const int32_t *__restrict__ count,
                                                 int32_t threshold, uint8_t *__restrict__ dirty_next,
                                                 uint8_t *__restrict__ tile_active, int32_t *__restrict__ labels_out)
{
	if ((int)blockIdx.x >= *count) {
		return;
	}
	const int tile = list[blockIdx.x], lane = threadIdx.x;
	int tx, ty, tz, ox, oy, oz;
	tile_origin(g, tile, tx, ty, tz, ox, oy, oz);

	__shared__ int32_t halo[kFaces];
	__shared__ int32_t fo[kFaces];
	__shared__ BfsShared bs;
	__shared__ uint32_t openw[K][R];
	__shared__ uint8_t marks[27];
	__shared__ int32_t lab[NV];

	uint32_t o[K], seed, excess;
	load_rows(g, bits, ntiles, tile, ox, oy, oz, lane, threshold, o, seed, excess);
	const uint32_t valid = valid_word(g, ox, oy, oz, lane);
	if (lane < R) {
#pragma unroll
		for (int d = 0; d < K; ++d) {
			openw[d][lane] = o[d];
		}
	}
	if (lane < 27) {
		marks[lane] = 0;
	}
	load_halo(g, halo, tx, ty, tz, lane);
	int32_t *own = g.faces + (size_t)tile * kFaces;
	int32_t old[(kFaces + 63) / 64];
#pragma unroll
	for (int j = 0; j < (kFaces + 63) / 64; ++j) {
		const int t = lane + 64 * j;
		old[j] = t < kFaces ? own[t] : kLabelInf;
	}
	__syncthreads();
	compute_exits(bs, halo, openw, lane);
	__syncthreads();
	uint32_t vis;
	tile_bfs(bs, nullptr, o, seed, valid, fo, labels_out ? lab : nullptr, 0, lane, vis);

#pragma unroll
	for (int j = 0; j < (kFaces + 63) / 64; ++j) {
		const int t = lane + 64 * j;
		if (t >= kFaces) {
			continue;
		}
		const int32_t nl = fo[t] < old[j] ? fo[t] : old[j];
		if (nl != old[j]) {
			own[t] = nl;
			if (halo[t] > nl + 1) {
				marks[dir_slot(face_dir(t))] = 1;
			}
		}
	}
	if (lane < R) {
		g.reach[(size_t)tile * R + lane] = vis;
	}
	if (labels_out) {
		const int x = lane & 31;
		for (int k = 0; k < kVoxPerLane; ++k) {
			const int r = slot_row(lane, k);
			const int gx = ox + x, gy = oy + row_y(r), gz = oz + row_z(r);
			if (inside(g, gx, gy, gz)) {
				labels_out[voxel_index(g, gx, gy, gz)] = lab[r * TX + x];
			}
		}
	}
	const bool active = __ballot(lane < R && (excess & vis) != 0u) != 0ull;
	__syncthreads();
	if (lane < 27 && marks[lane]) {
		mark_neighbour_tile(g, dirty_next, tx, ty, tz, lane);
	}
	if (lane == 0) {
		tile_active[tile] = active ? 1 : 0;
		atomicAdd(counter_slot(g, 1, tile), 1ull);
	}
}

// ---------------------------------------------------------------------------------------------
// Discharge one tile (colour phases: no neighbouring tile runs at the same time).
struct DischargeShared {
	int32_t halo[kFaces];
	int32_t fo[kFaces];
	BfsShared bs;
	LevelStore store;
	int32_t rl[K][NV];
	int32_t el[NV];
	int32_t rtl[NV];
	uint32_t dm[NV / 4];     // per-voxel byte of residual directions written in this visit
	uint16_t alist[NV];
	uint32_t openw[K][R], seedw[R], exw[R], validw[R];
	uint8_t marks[27];
};

__device__ inline void mark_dir(uint32_t *dm, int i, int d)
{
	atomicOr(&dm[i >> 2], (1u << d) << (8 * (i & 3)));
}

// Push the excess of active voxel i, at stored level index jj (level `level`), along admissible
// edges, into t first. jj - 1 indexes the stored level below (if any).
__device__ inline void push_voxel(DischargeShared &sh, int i, int jj, int32_t level, long long &sink)
{
	const int r = i >> 5, x = i & 31;
	const uint32_t bit = 1u << x;
	int32_t ex = sh.el[i];
	if (ex <= 0) {
		return;
	}
	if (sh.rtl[i] > 0) {
		const int32_t f = ex < sh.rtl[i] ? ex : sh.rtl[i];
		sh.rtl[i] -= f;
		ex -= f;
		sink += f;
		if (sh.rtl[i] == 0) {
			atomicAnd(&sh.seedw[r], ~bit);
		}
	}
	const bool below = jj > 0 && sh.store.lvl[jj - 1] == level - 1;
	int32_t rr[K];
	bool adm[K];
#pragma unroll
	for (int d = 0; d < K; ++d) {
		rr[d] = sh.rl[d][i];
		if (leaves(x, r, d)) {
			adm[d] = sh.halo[face_entry(d, r, x)] == level - 1;
		} else {
			const int j = neighbour_index(i, d);
			adm[d] = below && ((sh.store.ls[jj - 1][j >> 5] >> (j & 31)) & 1u);
		}
	}
#pragma unroll
	for (int d = 0; d < K; ++d) {
		if (ex <= 0 || rr[d] <= 0 || !adm[d]) {
			continue;
		}
		const int32_t f = ex < rr[d] ? ex : rr[d];
		ex -= f;
		rr[d] -= f;
		sh.rl[d][i] = rr[d];
		mark_dir(sh.dm, i, d);
		if (rr[d] == 0) {
			atomicAnd(&sh.openw[d][r], ~bit);
		}
		if (!leaves(x, r, d)) {
			const int j = neighbour_index(i, d);
			const int od = opposite<6>(d);
			const uint32_t jb = 1u << (j & 31);
			sh.rl[od][j] += f;  // only i pushes along this edge in this level
			mark_dir(sh.dm, j, od);
			atomicOr(&sh.openw[od][j >> 5], jb);
			atomicAdd(&sh.el[j], f);
			atomicOr(&sh.exw[j >> 5], jb);
		}
	}
	sh.el[i] = ex;
	if (ex == 0) {
		atomicAnd(&sh.exw[r], ~bit);
	}
}

__global__ __launch_bounds__(64) void k_wave_discharge(GridView g, uint32_t *__restrict__ bits, int ntiles,
                                                       const int32_t *__restrict__ list,
                                                       const int32_t *__restrict__ count, int max_iters,
                                                       uint8_t *__restrict__ tile_active)
{
	using D = Dirs<6>;
	if ((int)blockIdx.x >= *count) {
		return;
	}
	const int tile = list[blockIdx.x], lane = threadIdx.x;
	int tx, ty, tz, ox, oy, oz;
	tile_origin(g, tile, tx, ty, tz, ox, oy, oz);
	__shared__ DischargeShared sh;
	const bool row_lane = lane < R;
	const int x = lane & 31;

	// ---- load
	if (row_lane) {
#pragma unroll
		for (int d = 0; d < K; ++d) {
			sh.openw[d][lane] = bits[plane_word(d, tile, lane, ntiles)];
		}
		sh.seedw[lane] = bits[plane_word(kSeedPlane, tile, lane, ntiles)];
		sh.exw[lane] = bits[plane_word(kExcessPlane, tile, lane, ntiles)];
		sh.validw[lane] = valid_word(g, ox, oy, oz, lane);
	}
	if (lane < 27) {
		sh.marks[lane] = 0;
	}
	load_halo(g, sh.halo, tx, ty, tz, lane);
	int32_t e0[kVoxPerLane], rt0[kVoxPerLane];
	for (int k = 0; k < kVoxPerLane; ++k) {
		const int r = slot_row(lane, k);
		const int i = r * TX + x;
		const int gx = ox + x, gy = oy + row_y(r), gz = oz + row_z(r);
		int32_t ev = 0, rtv = 0;
		if (inside(g, gx, gy, gz)) {
			const int64_t v = voxel_index(g, gx, gy, gz);
			ev = g.e[v];
			rtv = g.rt[v];
#pragma unroll
			for (int d = 0; d < K; ++d) {
				sh.rl[d][i] = g.r[(int64_t)d * g.nvox + v];
			}
		} else {
#pragma unroll
			for (int d = 0; d < K; ++d) {
				sh.rl[d][i] = 0;
			}
		}
		sh.el[i] = ev;
		sh.rtl[i] = rtv;
		e0[k] = ev;
		rt0[k] = rtv;
	}
	for (int t = lane; t < NV / 4; t += 64) {
		sh.dm[t] = 0u;
	}
	__syncthreads();

	long long sink = 0;
	uint32_t vis = 0u;
	for (int it = 0; it < max_iters; ++it) {
		// ---- region relabel: exact distances inside the tile, halo labels fixed
		uint32_t o[K];
		const uint32_t valid = row_lane ? sh.validw[lane] : 0u;
		const uint32_t seed = row_lane ? sh.seedw[lane] : 0u;
#pragma unroll
		for (int d = 0; d < K; ++d) {
			o[d] = row_lane ? sh.openw[d][lane] : 0u;
		}
		compute_exits(sh.bs, sh.halo, sh.openw, lane);
		__syncthreads();
		const int levels = tile_bfs(sh.bs, &sh.store, o, seed, valid, sh.fo, nullptr, 0, lane, vis);
		if (__ballot(row_lane && (sh.exw[lane] & vis) != 0u) == 0ull) {
			break;
		}
		// ---- top-down push sweep, in chunks of kStore levels from the top
		int stored_lo = 0;  // the BFS above kept level indices [0, kStore]
		for (int hi = levels; hi > 0; hi -= kStore) {
			const int lo = hi - kStore > 0 ? hi - kStore : 0;
			const int want = lo > 0 ? lo - 1 : 0;  // keep the level below the chunk too
			if (want != stored_lo) {
				// same open words and sources as above, so the same levels
				uint32_t unused;
				tile_bfs(sh.bs, &sh.store, o, seed, valid, nullptr, nullptr, want, lane, unused);
				stored_lo = want;
			}
			for (int j = hi - 1; j >= lo; --j) {
				const int jj = j - stored_lo;
				const int32_t level = sh.store.lvl[jj];
				const uint32_t A = row_lane ? (sh.store.ls[jj][lane] & sh.exw[lane]) : 0u;
				int total;
				const int base = wave_exclusive_prefix(__popc(A), lane, total);
				if (total == 0) {
					continue;
				}
				{
					uint32_t b = A;
					int pos = base;
					while (b) {
						const int bb = __ffs((int)b) - 1;
						sh.alist[pos++] = (uint16_t)(lane * TX + bb);
						b &= b - 1u;
					}
				}
				__syncthreads();
				for (int t = lane; t < total; t += 64) {
					push_voxel(sh, sh.alist[t], jj, level, sink);
				}
				__syncthreads();
			}
		}
	}

	// ---- write back
	__syncthreads();
	int32_t *own = g.faces + (size_t)tile * kFaces;
	for (int t = lane; t < kFaces; t += 64) {
		own[t] = sh.fo[t];
	}
	for (int k = 0; k < kVoxPerLane; ++k) {
		const int r = slot_row(lane, k);
		const int i = r * TX + x;
		const int gx = ox + x, gy = oy + row_y(r), gz = oz + row_z(r);
		if (!inside(g, gx, gy, gz)) {
			continue;
		}
		const int64_t v = voxel_index(g, gx, gy, gz);
		if (sh.el[i] != e0[k]) {
			g.e[v] = sh.el[i];
		}
		if (sh.rtl[i] != rt0[k]) {
			g.rt[v] = sh.rtl[i];
		}
		const uint32_t dirty = (sh.dm[i >> 2] >> (8 * (i & 3))) & 0xffu;
		if (!dirty) {
			continue;
		}
#pragma unroll
		for (int d = 0; d < K; ++d) {
			if (!(dirty & (1u << d))) {
				continue;
			}
			const int32_t nr = sh.rl[d][i];
			if (leaves(x, r, d)) {
				const int ux = gx + D::dx(d), uy = gy + D::dy(d), uz = gz + D::dz(d);
				const int32_t old = g.r[(int64_t)d * g.nvox + v];
				const int32_t f = old - nr;
				if (f > 0 && inside(g, ux, uy, uz)) {
					// the halo voxel's tile is not running: this voxel is the only writer of the edge
					const int64_t u = voxel_index(g, ux, uy, uz);
					const int od = opposite<6>(d);
					atomicAdd(&g.e[u], f);
					g.r[(int64_t)od * g.nvox + u] += f;
					const int ntx = tx + D::dx(d), nty = ty + D::dy(d), ntz = tz + D::dz(d);
					const int ntile = (ntz * g.nty + nty) * g.ntx + ntx;
					const int nrow = (uz - ntz * TZ) * TY + (uy - nty * TY);
					const uint32_t nb = 1u << (ux - ntx * TX);
					atomicOr(&bits[plane_word(kExcessPlane, ntile, nrow, ntiles)], nb);
					atomicOr(&bits[plane_word(od, ntile, nrow, ntiles)], nb);
					sh.marks[dir_slot(d)] = 1;
				}
			}
			g.r[(int64_t)d * g.nvox + v] = nr;
		}
	}
	if (row_lane) {
#pragma unroll
		for (int d = 0; d < K; ++d) {
			bits[plane_word(d, tile, lane, ntiles)] = sh.openw[d][lane];
		}
		bits[plane_word(kSeedPlane, tile, lane, ntiles)] = sh.seedw[lane];
		bits[plane_word(kExcessPlane, tile, lane, ntiles)] = sh.exw[lane];
	}
	const bool active = __ballot(row_lane && (sh.exw[lane] & vis) != 0u) != 0ull;
	const long long total = wave_sum_i64(sink);
	__syncthreads();
	if (lane < 27 && sh.marks[lane]) {
		mark_neighbour_tile(g, tile_active, tx, ty, tz, lane);
	}
	if (lane == 0) {
		if (total) {
			atomicAdd(flow_slot(g, tile), (unsigned long long)total);
		}
		tile_active[tile] = active ? 1 : 0;
		atomicAdd(counter_slot(g, 0, tile), 1ull);
	}
}

// ---------------------------------------------------------------------------------------------
// S_min closure pass on bit planes: S grows along open edges inside the tile to a fixpoint, then
// spreads across the tile faces into the neighbouring tiles' words. Bits only rise, and every
// write is an atomicOr, so tiles of one pass may update each other's words.
__global__ __launch_bounds__(64) void k_wave_closure(GridView g, const uint32_t *__restrict__ bits,
                                                     uint32_t *__restrict__ sbits, int ntiles,
                                                     const int32_t *__restrict__ list,
                                                     const int32_t *__restrict__ count, int first_pass,
                                                     uint8_t *__restrict__ dirty_next)
{
	if ((int)blockIdx.x >= *count) {
		return;
	}
	const int tile = list[blockIdx.x], lane = threadIdx.x;
	int tx, ty, tz, ox, oy, oz;
	tile_origin(g, tile, tx, ty, tz, ox, oy, oz);
	const bool row_lane = lane < R;
	const int r = lane;
	uint32_t o[K] = {0, 0, 0, 0, 0, 0};
	uint32_t S = 0u, S0 = 0u;
	if (row_lane) {
#pragma unroll
		for (int d = 0; d < K; ++d) {
			o[d] = bits[plane_word(d, tile, r, ntiles)];
		}
		S0 = sbits[(size_t)tile * R + r];
		S = S0 | (first_pass ? bits[plane_word(kExcessPlane, tile, r, ntiles)] : 0u);
	}
	const int y = row_y(r), z = row_z(r);
	for (;;) {
		// S(u) |= S(v) for open v -> u: from x+1 (its -x edge), x-1 (+x), y+1 (-y), y-1 (+y),
		// z+1 (-z), z-1 (+z)
		const uint32_t from_yp = lane_read(S & o[1], r + 1), from_ym = lane_read(S & o[4], r - 1);
		const uint32_t from_zp = lane_read(S & o[2], r + TY), from_zm = lane_read(S & o[5], r - TY);
		uint32_t n = S | ((S & o[0]) >> 1) | ((S & o[3]) << 1);
		n |= y < TY - 1 ? from_yp : 0u;
		n |= y > 0 ? from_ym : 0u;
		n |= z < TZ - 1 ? from_zp : 0u;
		n |= z > 0 ? from_zm : 0u;
		if (!row_lane) {
			n = 0u;
		}
		const bool changed = __ballot(n != S) != 0ull;
		S = n;
		if (!changed) {
			break;
		}
	}
	__shared__ uint8_t marks[27];
	if (lane < 27) {
		marks[lane] = 0;
	}
	__syncthreads();
	if (row_lane) {
		if (S != S0) {
			atomicOr(&sbits[(size_t)tile * R + r], S & ~S0);
		}
		auto send = [&](int d, uint32_t w, int nrow) {
			if (!w) {
				return;
			}
			using D = Dirs<6>;
			const int ntx = tx + D::dx(d), nty = ty + D::dy(d), ntz = tz + D::dz(d);
			if (ntx < 0 || ntx >= g.ntx || nty < 0 || nty >= g.nty || ntz < 0 || ntz >= g.ntz) {
				return;
			}
			const int ntile = (ntz * g.nty + nty) * g.ntx + ntx;
			const uint32_t prev = atomicOr(&sbits[(size_t)ntile * R + nrow], w);
			if (w & ~prev) {
				marks[dir_slot(d)] = 1;
			}
		};
		send(0, ((S & o[0]) & 1u) << 31, r);
		send(3, ((S & o[3]) >> 31) & 1u, r);
		if (y == 0) {
			send(1, S & o[1], z * TY + (TY - 1));
		}
		if (y == TY - 1) {
			send(4, S & o[4], z * TY);
		}
		if (z == 0) {
			send(2, S & o[2], (TZ - 1) * TY + y);
		}
		if (z == TZ - 1) {
			send(5, S & o[5], y);
		}
	}
	__syncthreads();
	if (lane < 27 && marks[lane]) {
		mark_neighbour_tile(g, dirty_next, tx, ty, tz, lane);
	}
	if (lane == 0) {
		atomicAdd(counter_slot(g, 2, tile), 1ull);
	}
}

// One tile-major bit plane -> one byte per voxel in voxel order (set: `one`, clear: `zero`).
__global__ __launch_bounds__(64) void k_wave_plane_bytes(GridView g, const uint32_t *__restrict__ plane,
                                                         uint8_t *__restrict__ out, uint8_t one, uint8_t zero)
{
	const int tile = blockIdx.x, lane = threadIdx.x;
	int tx, ty, tz, ox, oy, oz;
	tile_origin(g, tile, tx, ty, tz, ox, oy, oz);
	const int x = lane & 31;
	for (int k = 0; k < kVoxPerLane; ++k) {
		const int r = slot_row(lane, k);
		const int gx = ox + x, gy = oy + row_y(r), gz = oz + row_z(r);
		if (inside(g, gx, gy, gz)) {
			out[voxel_index(g, gx, gy, gz)] = ((plane[(size_t)tile * R + r] >> x) & 1u) ? one : zero;
		}
	}
}

}  // namespace wave
}  // namespace cgp
