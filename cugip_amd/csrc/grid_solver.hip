// Host driver of the 3D grid push-relabel: buffers, tile geometry and the
// relabel / discharge / cut phase loop (MinCut::compute of the reference,
// include/cugip/advanced_operations/detail/graph_cut_implementation.hpp:122-198).
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <memory>
#include <mutex>

#include "grid_kernels.h"
#include "solver.h"
#include "tuned_policy.h"

namespace cgp {

namespace {

template <int kConn, int kTX, int kTY, int kTZ>
struct ShapeTag {
	static constexpr int conn = kConn, tx = kTX, ty = kTY, tz = kTZ;
};

// Tile shapes compiled into the library (connectivity, x, y, z).
#define CGP_GRID_SHAPES(F) \
	F(6, 32, 8, 4)         \
	F(6, 32, 8, 8)         \
	F(6, 64, 4, 4)         \
	F(6, 32, 4, 4)         \
	F(6, 64, 4, 8)         \
	F(6, 16, 8, 4)         \
	F(6, 16, 4, 4)         \
	F(6, 32, 2, 4)         \
	F(26, 32, 4, 4)        \
	F(26, 16, 8, 4)        \
	F(26, 16, 4, 4)        \
	F(26, 32, 4, 2)

template <typename Fn>
void dispatch_shape(int conn, int tx, int ty, int tz, Fn &&fn)
{
#define CGP_TRY_SHAPE(C, X, Y, Z)                          \
	if (conn == C && tx == X && ty == Y && tz == Z) {      \
		fn(ShapeTag<C, X, Y, Z>{});                        \
		return;                                            \
	}
	CGP_GRID_SHAPES(CGP_TRY_SHAPE)
#undef CGP_TRY_SHAPE
	fail(CGP_ERR_INVALID, "unsupported tile shape " + std::to_string(tx) + "x" + std::to_string(ty) + "x" +
	                          std::to_string(tz) + " for " + std::to_string(conn) + "-connectivity");
}

bool shape_supported(int conn, int tx, int ty, int tz)
{
#define CGP_MATCH_SHAPE(C, X, Y, Z) \
	if (conn == C && tx == X && ty == Y && tz == Z) return true;
	CGP_GRID_SHAPES(CGP_MATCH_SHAPE)
#undef CGP_MATCH_SHAPE
	return false;
}

int blocks_for(int64_t n, int threads, int cap = 4096)
{
	int64_t b = (n + threads - 1) / threads;
	if (b < 1) {
		b = 1;
	}
	return (int)std::min<int64_t>(b, cap);
}

double ms_since(std::chrono::steady_clock::time_point t)
{
	return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t).count();
}

}  // namespace

GridSolver::~GridSolver()
{
	for (hipEvent_t ev : mEvents) {
		(void)hipEventDestroy(ev);
	}
}

template <typename F>
void GridSolver::launch(Kern kind, int64_t voxels, double bytes_per_voxel, F &&fn)
{
	if (!mProfile) {
		fn();
		CGP_LAUNCH_CHECK();
		return;
	}
	const size_t ev = 2 * mPending.size();
	while (mEvents.size() < ev + 2) {
		hipEvent_t e = nullptr;
		CGP_HIP(hipEventCreate(&e));
		mEvents.push_back(e);
	}
	CGP_HIP(hipEventRecord(mEvents[ev], mStream));
	fn();
	CGP_LAUNCH_CHECK();
	CGP_HIP(hipEventRecord(mEvents[ev + 1], mStream));
	mPending.push_back(PendingTiming{kind, voxels, bytes_per_voxel, ev});
}

// Per-kernel times of a profiled solve, once its stream is idle; the tile passes' voxel visits
// come from the visit counters.
void GridSolver::finish_profile(const cgp_stats &stats)
{
	CGP_HIP(hipStreamSynchronize(mStream));
	for (const PendingTiming &t : mPending) {
		float ms = 0.0f;
		CGP_HIP(hipEventElapsedTime(&ms, mEvents[t.event], mEvents[t.event + 1]));
		KernelRecord &rec = mProf[t.kind];
		rec.launches += 1;
		rec.ms += ms;
		rec.bytes_per_voxel = t.bytes_per_voxel;
		if (t.voxels > 0) {
			rec.voxels += t.voxels;
		}
	}
	mPending.clear();
	const int64_t tile_voxels = (int64_t)mShape.tx * mShape.ty * mShape.tz;
	mProf[kDischarge].voxels = stats.discharge_tile_visits * tile_voxels;
	mProf[kBfs].voxels = stats.bfs_tile_visits * tile_voxels;
	mProf[kClosure].voxels = stats.cut_tile_visits * tile_voxels;
}

int GridSolver::profile(cgp_kernel_profile *out, int capacity) const
{
	static const char *names[kKernCount] = {"grid_init", "grid_bfs", "grid_discharge", "grid_source_closure",
	                                        "compact_tiles", "fill"};
	int n = 0;
	for (int k = 0; k < kKernCount; ++k) {
		if (mProf[k].launches == 0) {
			continue;
		}
		if (out && n < capacity) {
			cgp_kernel_profile &p = out[n];
			std::memset(&p, 0, sizeof(p));
			std::strncpy(p.name, names[k], sizeof(p.name) - 1);
			p.launches = mProf[k].launches;
			p.voxel_visits = mProf[k].voxels;
			p.device_ms = mProf[k].ms;
			p.bytes_per_voxel = mProf[k].bytes_per_voxel;
			p.algorithmic_bytes = (double)mProf[k].voxels * mProf[k].bytes_per_voxel;
		}
		++n;
	}
	return n;
}

GridSolver::Shape GridSolver::pick_shape(const cgp_policy &policy)
{
	if (policy.tile_x || policy.tile_y || policy.tile_z) {
		if (!shape_supported(mConn, policy.tile_x, policy.tile_y, policy.tile_z)) {
			fail(CGP_ERR_INVALID, "unsupported tile shape " + std::to_string(policy.tile_x) + "x" +
			                          std::to_string(policy.tile_y) + "x" + std::to_string(policy.tile_z));
		}
		return {policy.tile_x, policy.tile_y, policy.tile_z};
	}
	const TunedPolicy tp = tuned_policy(mConn, mInst.size() > 1, mVox / (int64_t)mInst.size());
	return Shape{tp.tile_x, tp.tile_y, tp.tile_z};
}

void GridSolver::set_grid(int nx, int ny, int nz, int conn, const int32_t *cs, const int32_t *ct, const int32_t *cn,
                          const int32_t *cnb, bool device_pointers)
{
	const GridSpec spec{nx, ny, nz, cs, ct, cn, cnb};
	set_batch(conn, &spec, 1, device_pointers);
}

std::vector<InstanceDesc> GridSolver::describe(int conn, const GridSpec *specs, int count)
{
	if (count <= 0 || !specs) {
		fail(CGP_ERR_INVALID, "a batch needs at least one grid");
	}
	if (conn != 6 && conn != 26) {
		fail(CGP_ERR_INVALID, "connectivity must be 6 or 26");
	}
	std::vector<InstanceDesc> inst((size_t)count);
	int64_t total = 0;
	for (int i = 0; i < count; ++i) {
		const GridSpec &sp = specs[i];
		const std::string which = count > 1 ? "grid " + std::to_string(i) + ": " : std::string();
		if (sp.nx <= 0 || sp.ny <= 0 || sp.nz <= 0) {
			fail(CGP_ERR_INVALID, which + "grid dimensions must be positive");
		}
		const int64_t vox = (int64_t)sp.nx * sp.ny * sp.nz;
		if (vox >= (int64_t)1 << 31) {
			fail(CGP_ERR_INVALID, which + "grids of 2^31 voxels or more are not supported");
		}
		if ((int64_t)sp.nx * sp.ny >= (int64_t)1 << 26) {
			// kernels address the voxels of a tile with 32-bit byte offsets from its first z-slice
			fail(CGP_ERR_INVALID, which + "z-slices of 2^26 voxels or more are not supported");
		}
		InstanceDesc &d = inst[(size_t)i];
		std::memset(&d, 0, sizeof(d));
		d.nx = sp.nx;
		d.ny = sp.ny;
		d.nz = sp.nz;
		d.conn = conn;
		d.vbase = total;
		d.nvox = vox;
		total += vox;
	}
	return inst;
}

void GridSolver::allocate_grids(int conn, const std::vector<InstanceDesc> &inst, bool has_back, const GridSpec *borrow)
{
	int64_t total = 0;
	for (const InstanceDesc &d : inst) {
		total += d.nvox;
	}
	const int half = conn / 2;
	const size_t count = inst.size();
	mSolved = false;
	mStarted = false;
	mVolumeKept = false;
	mTilesReady = false;
	mInst = inst;
	mFlows.assign(count, 0);
	mNx = inst[0].nx;
	mNy = inst[0].ny;
	mNz = inst[0].nz;
	mConn = conn;
	mVox = total;
	mHasBack = has_back;
	// until an init status says otherwise, the device-driven solve stages residuals as int16 (it
	// switches when the init kernel finds an edge whose capacity pair exceeds 32767)
	mWideResiduals = false;
	if (mArena) {
		mArena->used = 0;  // everything of the last grid is re-carved
	}
	if (borrow) {
		// the caller's capacity arrays, read in place (kernels never write them)
		mCs.borrow(const_cast<int32_t *>(borrow->cs), total);
		mCt.borrow(const_cast<int32_t *>(borrow->ct), total);
		mCn.borrow(const_cast<int32_t *>(borrow->cn), (size_t)total * half);
		mCnb.borrow(const_cast<int32_t *>(borrow->cnb), mHasBack ? (size_t)total * half : 0);
	} else {
		mCs.allocate(total, mArena);
		mCt.allocate(total, mArena);
		mCn.allocate((size_t)total * half, mArena);
		if (mHasBack) {
			mCnb.allocate((size_t)total * half, mArena);
		} else {
			mCnb.release();
		}
	}
	mE.allocate(total, mArena);
	mRt.allocate(total, mArena);
	mH.allocate(total, mArena);
	mR.allocate((size_t)total * conn, mArena);
	mS.allocate(total, mArena);
	mOmask.allocate(total, mArena);
	// the slotted visit counters (3 kinds) and then the per-grid flows, one region: one kernel sums
	// both, in the order of the totals of the read-back and of the control block
	mSlots.allocate((3 + count) * kSlotWords, mArena);
	mCounters.borrow(mSlots.get(), 3 * kSlotWords);
	mFlow.borrow(mSlots.get() + 3 * kSlotWords, count * kSlotWords);
	// one region read back whole by the closure's checks: the counts (int32 [0] list count, [1] [2]
	// count_two, [3] closure reached t), the init status (int32 [4]), then the 3 visit-counter totals
	// and the per-grid flow totals; every count read brings the status along
	mReadback.allocate(kReadbackHead + count, mArena);
	mCount.borrow(reinterpret_cast<int32_t *>(mReadback.get()), 4);
	mStatus.borrow(reinterpret_cast<int32_t *>(mReadback.get()) + 4, 1);
	mCounterTotal.borrow(mReadback.get() + 3, 3);
	mFlowTotal.borrow(mReadback.get() + kReadbackHead, count);
	// control block of the device-driven solve, then its visit-counter and per-grid flow totals
	mCtl.allocate(kCtlTotals + 3 + count, mArena);
	// read-backs: flows + visit counters, or the control block (pageable host memory in arena mode)
	static_assert(kCtlTotals >= 8, "the host scratch also holds the deferred counts after the read-back");
	mHostScratch.allocate(kCtlTotals + 3 + count, mArena != nullptr);
	++mGeneration;
	mRelabelCycles = 2;
	mRelabelFits = 0;
}

void GridSolver::set_batch(int conn, const GridSpec *specs, int count, bool device_pointers)
{
	const std::vector<InstanceDesc> inst = describe(conn, specs, count);
	bool any_back = false;
	for (int i = 0; i < count; ++i) {
		if (!specs[i].cs || !specs[i].ct || !specs[i].cn) {
			fail(CGP_ERR_INVALID, (count > 1 ? "grid " + std::to_string(i) + ": " : std::string()) +
			                          "capacity pointers must not be null");
		}
		any_back |= specs[i].cnb != nullptr;
	}
	allocate_grids(conn, inst, any_back);
	const int half = conn / 2;
	const hipMemcpyKind kind = device_pointers ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice;
	for (int i = 0; i < count; ++i) {
		const GridSpec &sp = specs[i];
		const InstanceDesc &d = inst[(size_t)i];
		const size_t vb = (size_t)d.nvox * sizeof(int32_t);
		CGP_HIP(hipMemcpyAsync(mCs.get() + d.vbase, sp.cs, vb, kind, mStream));
		CGP_HIP(hipMemcpyAsync(mCt.get() + d.vbase, sp.ct, vb, kind, mStream));
		CGP_HIP(hipMemcpyAsync(mCn.get() + d.vbase * half, sp.cn, vb * half, kind, mStream));
		if (mHasBack) {
			// a grid without back capacities is symmetric: its back planes are its forward planes
			CGP_HIP(hipMemcpyAsync(mCnb.get() + d.vbase * half, sp.cnb ? sp.cnb : sp.cn, vb * half, kind, mStream));
		}
	}
	CGP_HIP(hipStreamSynchronize(mStream));
}

void GridSolver::set_grid_in_place(int nx, int ny, int nz, int conn, const int32_t *cs, const int32_t *ct,
                                   const int32_t *cn, const int32_t *cnb)
{
	if (!cs || !ct || !cn) {
		fail(CGP_ERR_INVALID, "capacity pointers must not be null");
	}
	const GridSpec spec{nx, ny, nz, cs, ct, cn, cnb};
	const std::vector<InstanceDesc> inst = describe(conn, &spec, 1);
	allocate_grids(conn, inst, cnb != nullptr, &spec);
}

void GridSolver::set_grid_volume(int nx, int ny, int nz, int conn, const uint8_t *volume, const uint8_t *markers,
                                 const cgp_volume_weights &weights, bool device_pointers)
{
	if (!volume) {
		fail(CGP_ERR_INVALID, "volume pointer must not be null");
	}
	const GridSpec spec{nx, ny, nz, nullptr, nullptr, nullptr, nullptr};
	const std::vector<InstanceDesc> inst = describe(conn, &spec, 1);
	for (int i = 0; i < 256; ++i) {
		bool bad = weights.tlink_source[i] < 0 || weights.tlink_sink[i] < 0;
		for (int c = 0; c < 3; ++c) {
			bad |= weights.nlink[c][i] < 0;
		}
		if (bad) {
			fail(CGP_ERR_INVALID, "capacity tables must be non-negative");
		}
	}
	if (weights.marker_capacity < 0 || weights.border < 0) {
		fail(CGP_ERR_INVALID, "marker_capacity and border must be non-negative");
	}
	allocate_grids(conn, inst, false);
	const int64_t n = inst[0].nvox;
	DeviceBuffer<uint8_t> mark;
	DeviceBuffer<cgp_volume_weights> w(1);
	const uint8_t *dm = markers;
	// the volume stays on the device (the learned cut proposal reads it)
	mVolume.allocate(n);
	CGP_HIP(hipMemcpyAsync(mVolume.get(), volume, n, device_pointers ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice,
	                       mStream));
	const uint8_t *dv = mVolume.get();
	if (!device_pointers) {
		if (markers) {
			mark.allocate(n);
			CGP_HIP(hipMemcpyAsync(mark.get(), markers, n, hipMemcpyHostToDevice, mStream));
			dm = mark.get();
		}
	}
	CGP_HIP(hipMemcpyAsync(w.get(), &weights, sizeof(weights), hipMemcpyHostToDevice, mStream));
	const int blocks = blocks_for(n, 256, 8192);
	if (conn == 6) {
		hipLaunchKernelGGL(k_caps_from_volume<6>, dim3(blocks), dim3(256), 0, mStream, nx, ny, nz, dv, dm, w.get(),
		                   mCs.get(), mCt.get(), mCn.get());
	} else {
		hipLaunchKernelGGL(k_caps_from_volume<26>, dim3(blocks), dim3(256), 0, mStream, nx, ny, nz, dv, dm, w.get(),
		                   mCs.get(), mCt.get(), mCn.get());
	}
	CGP_LAUNCH_CHECK();
	CGP_HIP(hipStreamSynchronize(mStream));
	mVolumeKept = true;
}

int64_t GridSolver::cut_capacity(int index, const uint8_t *host_side)
{
	if (mInst.empty()) {
		fail(CGP_ERR_STATE, "no grid set");
	}
	const InstanceDesc &d = instance(index);
	DeviceBuffer<uint8_t> side(d.nvox);
	CGP_HIP(hipMemcpyAsync(side.get(), host_side, d.nvox, hipMemcpyHostToDevice, mStream));
	return cut_capacity_device(index, side.get());
}

int64_t GridSolver::cut_capacity_device(int index, const uint8_t *device_side)
{
	const InstanceDesc &d = instance(index);
	const int half = mConn / 2;
	const uint8_t *side = device_side;
	if (!mCutSlots.get()) {
		mCutSlots.allocate(kSlotWords + 1);  // kSlots slots, then their sum
	}
	unsigned long long *slots = mCutSlots.get();
	hipLaunchKernelGGL(k_fill_u8, dim3(blocks_for(8 * kSlotWords, 256)), dim3(256), 0, mStream,
	                   reinterpret_cast<uint8_t *>(slots), (int64_t)(8 * kSlotWords), (uint8_t)0);
	const int32_t *cn = mCn.get() + d.vbase * half;
	const int32_t *cnb = mHasBack ? mCnb.get() + d.vbase * half : nullptr;
	const int blocks = blocks_for(d.nvox, 256, 8192);
	if (mConn == 6) {
		hipLaunchKernelGGL(k_cut_capacity<6>, dim3(blocks), dim3(256), 0, mStream, d.nx, d.ny, d.nz, side,
		                   mCs.get() + d.vbase, mCt.get() + d.vbase, cn, cnb, slots);
	} else {
		hipLaunchKernelGGL(k_cut_capacity<26>, dim3(blocks), dim3(256), 0, mStream, d.nx, d.ny, d.nz, side,
		                   mCs.get() + d.vbase, mCt.get() + d.vbase, cn, cnb, slots);
	}
	CGP_LAUNCH_CHECK();
	hipLaunchKernelGGL(k_sum_slots, dim3(1), dim3(64), 0, mStream, slots, (int64_t)1, slots + kSlotWords);
	CGP_LAUNCH_CHECK();
	unsigned long long *pinned = reinterpret_cast<unsigned long long *>(mHostScratch.get());
	CGP_HIP(hipMemcpyAsync(pinned, slots + kSlotWords, sizeof(unsigned long long), hipMemcpyDeviceToHost, mStream));
	CGP_HIP(hipStreamSynchronize(mStream));
	return (int64_t)pinned[0];
}

void GridSolver::capacities(int index, int32_t *cs, int32_t *ct, int32_t *cn, int32_t *cnb)
{
	if (mInst.empty()) {
		fail(CGP_ERR_STATE, "no grid set");
	}
	const InstanceDesc &d = instance(index);
	const int half = mConn / 2;
	const size_t vb = (size_t)d.nvox * sizeof(int32_t);
	CGP_HIP(hipMemcpyAsync(cs, mCs.get() + d.vbase, vb, hipMemcpyDeviceToHost, mStream));
	CGP_HIP(hipMemcpyAsync(ct, mCt.get() + d.vbase, vb, hipMemcpyDeviceToHost, mStream));
	CGP_HIP(hipMemcpyAsync(cn, mCn.get() + d.vbase * half, vb * half, hipMemcpyDeviceToHost, mStream));
	const int32_t *back = mHasBack ? mCnb.get() : mCn.get();
	CGP_HIP(hipMemcpyAsync(cnb, back + d.vbase * half, vb * half, hipMemcpyDeviceToHost, mStream));
	CGP_HIP(hipStreamSynchronize(mStream));
}

const InstanceDesc &GridSolver::instance(int index) const
{
	if (index < 0 || index >= (int)mInst.size()) {
		fail(CGP_ERR_INVALID, "grid index " + std::to_string(index) + " out of range (batch of " +
		                          std::to_string(mInst.size()) + ")");
	}
	return mInst[(size_t)index];
}

// Colours of every tile inside its instance: face parity (tx + ty + tz) & 1, or with corners the
// parity per axis, so that no two tiles of a colour share a face, edge or corner.
void GridSolver::build_colours(ColourSet &set, int ncolours, bool corners)
{
	std::vector<uint8_t> colour((size_t)mTiles);
	std::vector<int> per_colour(ncolours, 0);
	for (const InstanceDesc &d : mInst) {
		const int n = d.ntx * d.nty * d.ntz;
		for (int t = 0; t < n; ++t) {
			const int tx = t % d.ntx, ty = (t / d.ntx) % d.nty, tz = t / (d.ntx * d.nty);
			const int c = corners ? ((tx & 1) | ((ty & 1) << 1) | ((tz & 1) << 2)) : ((tx + ty + tz) & 1);
			colour[(size_t)(d.tbase + t)] = (uint8_t)c;
			++per_colour[c];
		}
	}
	// largest colour class: the grid of a pass must cover every tile of one colour
	set.max_size = *std::max_element(per_colour.begin(), per_colour.end());
	set.colours = ncolours;
	set.colour.allocate(((size_t)mTiles + 31) / 32 * 32, mArena);
	CGP_HIP(hipMemcpyAsync(set.colour.get(), colour.data(), colour.size(), hipMemcpyHostToDevice, mStream));
	set.start.assign(ncolours, 0);
	set.size.assign(per_colour.begin(), per_colour.end());
	for (int c = 1; c < ncolours; ++c) {
		set.start[c] = set.start[c - 1] + set.size[c - 1];
	}
	std::vector<int32_t> grouped((size_t)mTiles + ncolours);
	std::vector<int> fill(set.start);
	for (int t = 0; t < mTiles; ++t) {
		grouped[(size_t)fill[colour[(size_t)t]]++] = t;
	}
	for (int c = 0; c < ncolours; ++c) {
		grouped[(size_t)mTiles + c] = set.size[c];
	}
	set.tiles.allocate(grouped.size(), mArena);
	CGP_HIP(hipMemcpyAsync(set.tiles.get(), grouped.data(), grouped.size() * sizeof(int32_t), hipMemcpyHostToDevice,
	                       mStream));
	CGP_HIP(hipStreamSynchronize(mStream));  // the host vectors go out of scope
}

void GridSolver::configure_tiles(const Shape &shape)
{
	if (mTilesReady && mReadyShape.tx == shape.tx && mReadyShape.ty == shape.ty && mReadyShape.tz == shape.tz) {
		return;
	}
	mTilesReady = false;
	++mGeneration;
	mShape = shape;
	mRelabelCycles = 2;
	mRelabelFits = 0;
	int64_t tiles = 0;
	for (InstanceDesc &d : mInst) {
		d.ntx = (d.nx + shape.tx - 1) / shape.tx;
		d.nty = (d.ny + shape.ty - 1) / shape.ty;
		d.ntz = (d.nz + shape.tz - 1) / shape.tz;
		d.tbase = (int32_t)tiles;
		tiles += (int64_t)d.ntx * d.nty * d.ntz;
		if (tiles >= (int64_t)1 << 31) {
			fail(CGP_ERR_INVALID, "a batch may hold at most 2^31 - 1 tiles");
		}
	}
	mNtx = mInst[0].ntx;
	mNty = mInst[0].nty;
	mNtz = mInst[0].ntz;
	mTiles = (int)tiles;
	std::vector<int32_t> owner(mInst.size() > 1 ? (size_t)mTiles : 0);
	for (size_t i = 0; i < mInst.size(); ++i) {
		const InstanceDesc &d = mInst[i];
		const int n = d.ntx * d.nty * d.ntz;
		for (int t = 0; t < n && !owner.empty(); ++t) {
			owner[(size_t)(d.tbase + t)] = (int32_t)i;
		}
	}
	// flag arrays padded to whole 32-byte chunks: the one-workgroup compaction reads them 16 bytes at a time
	const int64_t padded = ((int64_t)mTiles + 31) / 32 * 32;
	mActive.allocate(padded, mArena);
	mTileWork.allocate(padded, mArena);
	mDirtyA.allocate(padded, mArena);
	mDirtyB.allocate(padded, mArena);
	mCursors.allocate(kCursorWords, mArena);
	hipLaunchKernelGGL(k_fill_i32, dim3(1), dim3(64), 0, mStream, mCursors.get(), (int64_t)kCursorWords, 0,
	                   (const SolveControl *)nullptr, 0u);
	mDirtyAClean = mDirtyBClean = false;  // new flag arrays
	mAllTiles.allocate(mTiles + 1, mArena);
	std::vector<int32_t> all(mTiles + 1);
	for (int t = 0; t < mTiles; ++t) {
		all[t] = t;
	}
	all[mTiles] = mTiles;  // count slot for the all-tiles list
	CGP_HIP(hipMemcpyAsync(mAllTiles.get(), all.data(), all.size() * sizeof(int32_t), hipMemcpyHostToDevice, mStream));
	build_colours(mPhase, mConn == 6 ? 2 : 8, mConn != 6);
	build_colours(mSweep, 2, false);
	// beyond one workgroup's compaction, ordered discharge lists are laid out in class regions of
	// the largest colour class each
	mClassStride = mPhase.max_size;
	mList.allocate(mTiles > kCompactBlockTiles ? std::max<int64_t>(mTiles, (int64_t)kWorkClasses * mClassStride)
	                                           : (int64_t)mTiles,
	               mArena);
	// workgroups of the voxel-parallel init (init_blocks per grid)
	{
		std::vector<int32_t> first(mInst.size());
		int64_t blocks = 0;
		for (size_t i = 0; i < mInst.size(); ++i) {
			first[i] = (int32_t)blocks;
			blocks += init_blocks(mInst[i].nx, mInst[i].ny, mInst[i].nz, mConn);
		}
		mInitBlocks = (int)blocks;
		if (mInst.size() > 1) {
			std::vector<int32_t> owner_block((size_t)blocks);
			for (size_t i = 0; i < mInst.size(); ++i) {
				const int64_t end = i + 1 < mInst.size() ? first[i + 1] : blocks;
				for (int64_t b = first[i]; b < end; ++b) {
					owner_block[(size_t)b] = (int32_t)i;
				}
			}
			mInitBlockInst.allocate((size_t)blocks, mArena);
			mInitFirstBlock.allocate(mInst.size(), mArena);
			CGP_HIP(hipMemcpyAsync(mInitBlockInst.get(), owner_block.data(), owner_block.size() * sizeof(int32_t),
			                       hipMemcpyHostToDevice, mStream));
			CGP_HIP(hipMemcpyAsync(mInitFirstBlock.get(), first.data(), first.size() * sizeof(int32_t),
			                       hipMemcpyHostToDevice, mStream));
			CGP_HIP(hipStreamSynchronize(mStream));
		}
	}
	if (mInst.size() > 1) {
		mInstDev.allocate(mInst.size(), mArena);
		mTileInst.allocate(mTiles, mArena);
		CGP_HIP(hipMemcpyAsync(mInstDev.get(), mInst.data(), mInst.size() * sizeof(InstanceDesc),
		                       hipMemcpyHostToDevice, mStream));
		CGP_HIP(hipMemcpyAsync(mTileInst.get(), owner.data(), owner.size() * sizeof(int32_t), hipMemcpyHostToDevice,
		                       mStream));
	}
	CGP_HIP(hipStreamSynchronize(mStream));
	mTilesReady = true;
	mReadyShape = shape;
}

static GridView make_view(int nx, int ny, int nz, int ntx, int nty, int ntz, int64_t vox, DeviceBuffer<int32_t> &e,
                          DeviceBuffer<int32_t> &rt, DeviceBuffer<int32_t> &h, DeviceBuffer<int32_t> &r,
                          DeviceBuffer<uint8_t> &s, DeviceBuffer<uint32_t> &omask, DeviceBuffer<int32_t> &cs,
                          DeviceBuffer<int32_t> &ct, DeviceBuffer<int32_t> &cn, DeviceBuffer<int32_t> &cnb,
                          DeviceBuffer<unsigned long long> &flow, DeviceBuffer<int32_t> &status,
                          DeviceBuffer<unsigned long long> &counters)
{
	GridView g;
	g.nx = nx;
	g.ny = ny;
	g.nz = nz;
	g.ntx = ntx;
	g.nty = nty;
	g.ntz = ntz;
	g.nvox = vox;
	g.e = e.get();
	g.rt = rt.get();
	g.h = h.get();
	g.r = r.get();
	g.s = s.get();
	g.omask = omask.get();
	g.cs = cs.get();
	g.ct = ct.get();
	g.cn = cn.get();
	g.cnb = cnb.get();
	g.flow = flow.get();
	g.status = status.get();
	g.counters = counters.get();
	g.inst = nullptr;
	g.tile_inst = nullptr;
	g.tbase = 0;
	return g;
}

#define CGP_VIEW() batch_view(make_view(mNx, mNy, mNz, mNtx, mNty, mNtz, mVox, mE, mRt, mH, mR, mS, mOmask, mCs, \
                                       mCt, mCn, mCnb, mFlow, mStatus, mCounters),                             \
                             mInst.size() > 1 ? mInstDev.get() : nullptr, mTileInst.get())

// A batch's kernels find each tile's instance through the instance table.
static GridView batch_view(GridView g, const InstanceDesc *inst, const int32_t *tile_inst)
{
	if (inst) {
		g.inst = inst;
		g.tile_inst = tile_inst;
	}
	return g;
}

void GridSolver::launch_init(const SolveControl *ctl)
{
	GridView g = CGP_VIEW();
	// reads cs, ct and the n-link planes, writes e, rt, the K residual planes and the state word
	const double bytes = 4.0 * (2 + (mConn / 2) * (mHasBack ? 2 : 1)) + 4.0 * (3 + mConn);
	launch(kInit, mVox, bytes, [&] {
		const bool batch = mInst.size() > 1;
		const int32_t *block_inst = batch ? mInitBlockInst.get() : nullptr;
		const int32_t *first = batch ? mInitFirstBlock.get() : nullptr;
		if (mConn == 6) {
			hipLaunchKernelGGL(k_grid_init<6>, dim3(mInitBlocks), dim3(kInitThreads), 0, mStream, g, block_inst, first,
			                   ctl);
		} else {
			hipLaunchKernelGGL(k_grid_init<26>, dim3(mInitBlocks), dim3(kInitThreads), 0, mStream, g, block_inst,
			                   first, ctl);
		}
	});
}

void GridSolver::launch_compact(const ColourSet &set, uint8_t *flags, int colour, bool clear, const uint8_t *work,
                                const SolveControl *ctl, unsigned gate)
{
	if (mTiles <= kCompactBlockTiles) {
		launch(kCompact, 0, 0.0, [&] {
			hipLaunchKernelGGL(k_compact_block<1>, dim3(1), dim3(kCompactThreads), 0, mStream, flags, mTiles,
			                   set.colour.get(), colour, clear ? 1 : 0, mList.get(), mCount.get(),
			                   (const uint8_t *)nullptr, (int32_t *)nullptr, work, ctl, gate);
		});
		mListCount = mCount.get();
		mDischargeStride = 0;
		return;
	}
	const int blocks = (mTiles + kCompactTilesPerGroup - 1) / kCompactTilesPerGroup;
	if (work) {
		// ordered by work class into class regions (the lists of discharge phases never clear flags);
		// the cursors alternate between two sets, each compaction clearing the other
		int32_t *cursors = mCursors.get() + kWorkClasses * mOrderParity;
		int32_t *next = mCursors.get() + kWorkClasses * (mOrderParity ^ 1);
		launch(kCompact, 0, 0.0, [&] {
			const int oblocks = (mTiles + kOrderThreads * kCompactTilesPerThread - 1) / (kOrderThreads * kCompactTilesPerThread);
			hipLaunchKernelGGL(k_order_regions, dim3(oblocks), dim3(kOrderThreads), 0, mStream, (const uint8_t *)flags,
			                   mTiles, (const uint8_t *)set.colour.get(), colour, work, mList.get(), mClassStride,
			                   cursors, next, ctl, gate);
		});
		mListCount = cursors;
		mDischargeStride = mClassStride;
		mOrderParity ^= 1;
		return;
	}
	// the count alternates between two words, each compaction clearing the other
	int32_t *count = mCursors.get() + 2 * kWorkClasses + mListParity;
	int32_t *next = mCursors.get() + 2 * kWorkClasses + (mListParity ^ 1);
	launch(kCompact, 0, 0.0, [&] {
		hipLaunchKernelGGL(k_compact_tiles, dim3(blocks), dim3(kCompactThreads), 0, mStream, flags, mTiles,
		                   set.colour.get(), colour, clear ? 1 : 0, mList.get(), count, next, ctl, gate);
	});
	mListCount = count;
	mDischargeStride = 0;
	mListParity ^= 1;
}

void GridSolver::apply_status(int32_t status)
{
	mStatusPending = false;
	if (status & 1) {
		fail(CGP_ERR_INVALID, "capacities must be non-negative");
	}
	if (status & 2) {
		fail(CGP_ERR_OVERFLOW, "each voxel's source capacity plus incoming capacities, and each edge's two "
		                       "capacities, must sum below 2^31");
	}
	mWideResiduals = (status & 4) != 0;
}

// Counts the set flags of two arrays with one host read (which also brings a pending init status).
// The host copy holds the count words at int32 [1] [2] (and [3], the closure's reached flag); during
// a closure (mPrefetchTotals) the read also brings the visit-counter and flow totals, since the
// solve ends at one of the closure's checks (mTotalsRead then says the copy is current).
void GridSolver::count_two(uint8_t *a, uint8_t *b, int &na, int &nb, bool defer)
{
	int32_t *pinned = reinterpret_cast<int32_t *>(mHostScratch.get());
	const size_t n = mInst.size();
	if (mTiles <= kCompactBlockTiles) {
		launch(kCompact, 0, 0.0, [&] {
			hipLaunchKernelGGL(k_compact_block<1>, dim3(1), dim3(kCompactThreads), 0, mStream, a, mTiles,
			                   mSweep.colour.get(), -1, 0, (int32_t *)nullptr, mCount.get() + 1, (const uint8_t *)b,
			                   mCount.get() + 2, (const uint8_t *)nullptr, (const SolveControl *)nullptr, 0u);
		});
	} else {
		// with the totals prefetched, the count's last workgroup also sums the slots
		launch(kCompact, 0, 0.0, [&] {
			const int blocks = (mTiles + kCompactTilesPerGroup - 1) / kCompactTilesPerGroup;
			hipLaunchKernelGGL(k_count_flags2, dim3(blocks), dim3(1024), 0, mStream, (const uint8_t *)a,
			                   (const uint8_t *)b, mTiles, mCount.get() + 1, mCursors.get() + kCountAcc,
			                   mPrefetchTotals ? (const unsigned long long *)mSlots.get() : nullptr, (int64_t)(3 + n),
			                   mCounterTotal.get());
		});
	}
	size_t bytes = 6 * sizeof(int32_t);  // the counts and the init status
	if (mPrefetchTotals) {
		if (mTiles <= kCompactBlockTiles) {
			hipLaunchKernelGGL(k_sum_slots, dim3(blocks_for((int64_t)(3 + n), 256)), dim3(256), 0, mStream,
			                   mSlots.get(), (int64_t)(3 + n), mCounterTotal.get());
			CGP_LAUNCH_CHECK();
		}
		bytes = (kReadbackHead + n) * sizeof(unsigned long long);
	}
	if (defer) {
		// the counts go to their own slot: the next host wait reads them (deferred_counts)
		CGP_HIP(hipMemcpyAsync(const_cast<int32_t *>(deferred_counts()), mCount.get(), 4 * sizeof(int32_t),
		                       hipMemcpyDeviceToHost, mStream));
		mCheckDeferred = true;
		return;
	}
	CGP_HIP(hipMemcpyAsync(pinned, mCount.get(), bytes, hipMemcpyDeviceToHost, mStream));
	count_two_wait(na, nb, b != nullptr);
}

// Host copy of the counts of a deferred count_two (valid after the next host wait): [1] the first
// array's flags, [2] the second's.
const int32_t *GridSolver::deferred_counts() const
{
	return reinterpret_cast<const int32_t *>(mHostScratch.get()) + 2 * (kReadbackHead + mInst.size()) + 4;
}

// The host wait of a count_two: the counts (and a pending init status) once the stream got there.
void GridSolver::count_two_wait(int &na, int &nb, bool has_b)
{
	const int32_t *pinned = reinterpret_cast<const int32_t *>(mHostScratch.get());
	CGP_HIP(hipStreamSynchronize(mStream));
	mCheckDeferred = false;
	mTotalsRead = mPrefetchTotals;
	if (mStatusPending) {
		apply_status(pinned[4]);
	}
	na = pinned[1];
	nb = has_b ? pinned[2] : 0;
}

// Flow of every grid (into mFlows) and the three visit counters from the totals the last closure
// check read back.
int64_t GridSolver::totals_from_readback(unsigned long long *visits)
{
	const unsigned long long *words = reinterpret_cast<const unsigned long long *>(mHostScratch.get());
	for (int k = 0; k < 3; ++k) {
		visits[k] = words[3 + k];
	}
	int64_t total = 0;
	for (size_t i = 0; i < mInst.size(); ++i) {
		mFlows[i] = (int64_t)words[kReadbackHead + i];
		total += mFlows[i];
	}
	return total;
}

// Visits every tile once, colour by colour, then keeps visiting the dirty tiles colour by colour
// until a check finds none. Tiles of one colour never touch, so each pass reads the final state
// of the neighbours it depends on; the passes between two checks are chained on the device (each
// reads its tile count from device memory, blocks beyond it exit), so a relabel costs one or two
// host reads instead of one per pass. pass(list, count, grid, first) launches one pass.
// Returns the number of pass launches; `also` receives the flag count of also_count (if given)
// at the final check.
template <typename Pass>
int GridSolver::colour_sweeps(uint8_t *dirty, uint8_t *also_count, int &also, Pass &&pass, bool *stop, bool defer,
                              int first_cycles)
{
	constexpr int kCyclesPerCheck = 2;
	int launches = 0;
	mSweepChecks = 0;
	mSweepCycles = 0;
	const ColourSet &cs = mSweep;
	for (int c = 0; c < cs.colours; ++c) {
		if (cs.size[c] == 0) {
			continue;  // e.g. a grid one tile thick has no tile of the odd colours
		}
		pass(cs.tiles.get() + cs.start[c], cs.tiles.get() + mTiles + c, cs.size[c], c + 1);
		++launches;
		if (stop && mConn != 6 && c + 1 < cs.colours) {
			// a test that fails usually fails in its first pass; 26C solves take several rounds, so
			// their tests fail often: read the flag after each colour
			int32_t *pinned = reinterpret_cast<int32_t *>(mHostScratch.get());
			CGP_HIP(hipMemcpyAsync(pinned, mCount.get() + 3, sizeof(int32_t), hipMemcpyDeviceToHost, mStream));
			CGP_HIP(hipStreamSynchronize(mStream));
			if (pinned[0]) {
				*stop = true;
				return launches;
			}
		}
	}
	// the first sweep always leaves dirty tiles (the first colour ran before its neighbours), so the
	// first cycles (first_cycles of them) follow it without a check (one host wait fewer)
	for (bool check = false;; check = true) {
		if (check) {
			int pending = 0;
			++mSweepChecks;
			count_two(dirty, also_count, pending, also, defer);
			if (defer) {
				return launches;  // the caller queues more work, then waits for the counts
			}
			if (stop) {
				// the sweeps' own flag (the closure reaching t), read with the counts
				const int32_t *pinned = reinterpret_cast<const int32_t *>(mHostScratch.get());
				if (pinned[3]) {
					*stop = true;
					return launches;
				}
			}
			if (pending == 0) {
				return launches;
			}
		}
		const int cycles = check ? kCyclesPerCheck : first_cycles;
		for (int cycle = 0; cycle < cycles; ++cycle) {
			for (int c = 0; c < cs.colours; ++c) {
				launch_compact(cs, dirty, c, true);
				pass(mList.get(), mListCount, cs.max_size, 0);
				++launches;
			}
		}
		mSweepCycles += cycles;
	}
}

// One global-relabel pass over `grid` blocks taking tiles from list (count tiles).
void GridSolver::launch_bfs_pass(const int32_t *list, const int32_t *count, int grid, int32_t threshold,
                                 int64_t voxels, int first_mode)
{
	GridView g = CGP_VIEW();
	// exact passes read the state word and h and write h; threshold passes read r[K], rt, e, h
	const double bytes = threshold == 0 ? 12.0 : 4.0 * (mConn + 4);
	launch(kBfs, voxels, bytes, [&] {
		dispatch_shape(mConn, mShape.tx, mShape.ty, mShape.tz, [&](auto tag) {
			using S = decltype(tag);
			if constexpr (S::conn == 6 && S::tx * S::ty == 64) {
				// one-wave 6C tiles: the tile in registers
				hipLaunchKernelGGL((k_grid_bfs_wave<S::tx, S::ty, S::tz>),
				                   dim3((grid + kPassTilesPerBlock - 1) / kPassTilesPerBlock), dim3(64 * kPassTilesPerBlock),
				                   0, mStream, g, list,
				                   count, threshold, first_mode, mDirtyA.get(), mActive.get(), mTileWork.get());
				return;
			}
			hipLaunchKernelGGL((k_grid_bfs<S::conn, S::tx, S::ty, S::tz>), dim3(grid), dim3(S::tx * S::ty), 0,
			                   mStream, g, list, count, threshold, first_mode, mDirtyA.get(), mActive.get(),
			                   mTileWork.get());
		});
	});
}

// One S_min closure pass (first: seeded from the excess). reached (may be null) is set when the
// closure takes in a voxel with a residual sink link.
void GridSolver::launch_closure_pass(const int32_t *list, const int32_t *count, int grid, bool first,
                                     int32_t *reached, const SolveControl *ctl, int64_t voxels)
{
	GridView g = CGP_VIEW();
	// reads the state word and s, writes s
	launch(kClosure, voxels, 6.0, [&] {
		dispatch_shape(mConn, mShape.tx, mShape.ty, mShape.tz, [&](auto tag) {
			using S = decltype(tag);
			if constexpr (S::conn == 6 && S::tx * S::ty == 64) {
				// one-wave 6C tiles: the tile in registers
				hipLaunchKernelGGL((k_grid_closure_wave<S::tx, S::ty, S::tz>),
				                   dim3((grid + kPassTilesPerBlock - 1) / kPassTilesPerBlock), dim3(64 * kPassTilesPerBlock),
				                   0, mStream, g, list,
				                   count, first ? 1 : 0, mDirtyB.get(), reached, ctl);
				return;
			}
			hipLaunchKernelGGL((k_grid_source_closure<S::conn, S::tx, S::ty, S::tz>), dim3(grid), dim3(S::tx * S::ty),
			                   0, mStream, g, list, count, first ? 1 : 0, mDirtyB.get(), reached, ctl);
		});
	});
}

// One discharge phase over the compacted list (every tile of one colour at most; blocks beyond
// the count exit at once).
void GridSolver::launch_discharge(int max_iters, bool wide)
{
	GridView g = CGP_VIEW();
	const int blocks = mPhase.max_size;
	// reads and writes e, rt, h and r[K], writes the state word; one-wave 6C tiles keep the tile in
	// registers, the others stage it in LDS with int16 residuals unless some edge's capacity pair
	// exceeds 32767
	launch(kDischarge, -1, 8.0 * (mConn + 3) + 4.0, [&] {
		dispatch_shape(mConn, mShape.tx, mShape.ty, mShape.tz, [&](auto tag) {
			using S = decltype(tag);
			if constexpr (S::conn == 6 && S::tx * S::ty == 64 && S::tz <= 4) {
				// one-wave 6C tiles: the whole tile state in registers (int32 residuals, any width)
				hipLaunchKernelGGL((k_grid_discharge_wave<S::tx, S::ty, S::tz>),
				                   dim3((blocks + kDischargeTilesPerBlock - 1) / kDischargeTilesPerBlock),
				                   dim3(64 * kDischargeTilesPerBlock),
				                   0, mStream, g,
				                   mList.get(), mListCount, mDischargeStride, max_iters, mActive.get());
				return;
			}
			if (wide) {
				hipLaunchKernelGGL((k_grid_discharge<S::conn, S::tx, S::ty, S::tz, int32_t>), dim3(blocks),
				                   dim3(S::tx * S::ty), 0, mStream, g, mList.get(), mListCount, mDischargeStride,
				                   max_iters, mActive.get());
			} else {
				hipLaunchKernelGGL((k_grid_discharge<S::conn, S::tx, S::ty, S::tz, int16_t>), dim3(blocks),
				                   dim3(S::tx * S::ty), 0, mStream, g, mList.get(), mListCount, mDischargeStride,
				                   max_iters, mActive.get());
			}
		});
	});
}

int GridSolver::run_global_relabel(int32_t threshold, cgp_stats &stats, bool defer)
{
	const int64_t tile_voxels = (int64_t)mShape.tx * mShape.ty * mShape.tz;
	// 6-connected first sweeps start from kLabelInf themselves (first_mode); 26-connected tiles read
	// diagonal neighbours of their own sweep colour, so the labels are filled first
	if (mConn != 6) {
		launch(kFill, mVox, 4.0, [&] {
			hipLaunchKernelGGL(k_fill_i32, dim3(blocks_for(mVox, 256, 8192)), dim3(256), 0, mStream, mH.get(), mVox,
			                   kLabelInf, (const SolveControl *)nullptr, 0u);
		});
	}
	// a relabel that ran to its end left no dirty flag
	if (!mDirtyAClean) {
		launch(kFill, 0, 0.0, [&] {
			hipLaunchKernelGGL(k_fill_u8, dim3(blocks_for(mTiles, 256)), dim3(256), 0, mStream, mDirtyA.get(),
			                   (int64_t)mTiles, (uint8_t)0, (const SolveControl *)nullptr, 0u);
		});
	}
	mDirtyAClean = false;
	int active = 0;
	stats.bfs_passes += colour_sweeps(
	    mDirtyA.get(), mActive.get(), active,
	    [&](const int32_t *list, const int32_t *count, int grid, int first) {
		    launch_bfs_pass(list, count, grid, threshold, first ? grid * tile_voxels : -1, mConn == 6 ? first : 0);
	    },
	    nullptr, defer, mRelabelCycles);
	if (!mCheckDeferred) {
		learn_relabel_cycles(mSweepChecks > 1 ? mSweepCycles : 0);
	}
	mDirtyAClean = !mCheckDeferred;  // a deferred check may still find dirty tiles
	return active;
}

// Feedback from a relabel's checks: `needed` colour cycles when it ran past its first check (the
// next relabels check that late), 0 when it ended at its first check.
void GridSolver::learn_relabel_cycles(int needed)
{
	if (needed > 0) {
		mRelabelCycles = std::min(needed, (int)kMaxRelabelCycles);
		mRelabelFits = 0;
		return;
	}
	if (mRelabelCycles > 2 && ++mRelabelFits >= 8) {
		--mRelabelCycles;
		mRelabelFits = 0;
	}
}

void GridSolver::run_discharge_phase(int colour, int max_iters)
{
	// heaviest tiles first (work estimates of the last relabel)
	launch_compact(mPhase, mActive.get(), colour, false, mTileWork.get());
	launch_discharge(max_iters, mWideResiduals);
}

bool GridSolver::run_source_closure(cgp_stats &stats, bool check)
{
	const int64_t tile_voxels = (int64_t)mShape.tx * mShape.ty * mShape.tz;
	int32_t *reached = check ? mCount.get() + 3 : nullptr;
	// s, the reached flag and (unless the last closure consumed them all) the closure's tile flags
	launch(kFill, mVox, 1.0, [&] {
		hipLaunchKernelGGL(k_closure_reset, dim3(blocks_for(mVox / 16 + 1, 256, 4096)), dim3(256), 0, mStream, mS.get(),
		                   mVox, mCount.get() + 3, mDirtyBClean ? (uint8_t *)nullptr : mDirtyB.get(), (int64_t)mTiles,
		                   (const SolveControl *)nullptr, 0u);
	});
	mDirtyBClean = false;
	int unused = 0;
	bool hit = false;
	mTotalsRead = false;
	mPrefetchTotals = true;
	try {
		stats.cut_passes += colour_sweeps(
		    mDirtyB.get(), nullptr, unused,
		    [&](const int32_t *list, const int32_t *count, int grid, int first) {
			    launch_closure_pass(list, count, grid, first != 0, reached, nullptr, first ? grid * tile_voxels : -1);
		    },
		    check ? &hit : nullptr);
	} catch (...) {
		mPrefetchTotals = false;
		throw;
	}
	mPrefetchTotals = false;
	mDirtyBClean = !hit;  // the sweeps ran until no flag was left
	// (the last count of the sweeps read the flag after their last pass)
	return hit;
}

// Flow of every grid (summed into mFlows) and, with visits set, the three visit counters, with one
// host wait.
int64_t GridSolver::read_flow(unsigned long long *visits)
{
	const size_t n = mInst.size();
	unsigned long long *pinned = reinterpret_cast<unsigned long long *>(mHostScratch.get());
	hipLaunchKernelGGL(k_sum_slots, dim3(blocks_for((int64_t)(3 + n), 256)), dim3(256), 0, mStream, mSlots.get(),
	                   (int64_t)(3 + n), mCounterTotal.get());
	CGP_LAUNCH_CHECK();
	CGP_HIP(hipMemcpyAsync(pinned, mCounterTotal.get(), (3 + n) * sizeof(unsigned long long), hipMemcpyDeviceToHost,
	                       mStream));
	CGP_HIP(hipStreamSynchronize(mStream));
	if (visits) {
		visits[0] = pinned[0];
		visits[1] = pinned[1];
		visits[2] = pinned[2];
	}
	int64_t total = 0;
	for (size_t i = 0; i < n; ++i) {
		mFlows[i] = (int64_t)pinned[3 + i];
		total += mFlows[i];
	}
	return total;
}

cgp_policy default_policy()
{
	cgp_policy p;
	std::memset(&p, 0, sizeof(p));
	return p;
}

// ---- device-driven solve ------------------------------------------------------------------------

GridSolver::RoundGraph::~RoundGraph() { reset(); }

void GridSolver::RoundGraph::reset()
{
	if (exec) {
		(void)hipGraphExecDestroy(exec);
	}
	exec = nullptr;
}

// Records init plus one round of stages (relabel, discharge phases, termination closure or final
// closure), each gated by the control block, then the flow and visit-counter sums. Replayed until
// the control block says done: a replay whose stage is not the current one costs only launches of
// kernels that exit at once.
void GridSolver::capture_round_graph(int phases, int max_iters, bool wide)
{
	mGraph.reset();
	SolveControl *ctl = control();
	const int64_t tile_voxels = (int64_t)mShape.tx * mShape.ty * mShape.tz;
	const int ncolours = mConn == 6 ? 2 : 8;
	const int32_t *sweep_sizes = mSweep.tiles.get() + mTiles;
	unsigned long long *totals = mCtl.get() + kCtlTotals;
	const size_t n = mInst.size();
	hipGraph_t graph = nullptr;
	CGP_HIP(hipStreamBeginCapture(mStream, hipStreamCaptureModeThreadLocal));
	try {
		mCapturing = true;
		// init (gated: it runs once per solve; the status, flows and counters were cleared before the
		// first replay)
		launch_init(ctl);
		hipLaunchKernelGGL(k_ctl_after_init, dim3(1), dim3(64), 0, mStream, ctl, (const int32_t *)mStatus.get(),
		                   wide ? 1 : 0);
		// global relabel: from scratch, or continued when the last replay left dirty tiles
		hipLaunchKernelGGL(k_ctl_stage_begin, dim3(1), dim3(64), 0, mStream, ctl, kGateRelabelNew, sweep_sizes);
		if (mConn != 6) {
			hipLaunchKernelGGL(k_fill_i32, dim3(blocks_for(mVox, 256, 8192)), dim3(256), 0, mStream, mH.get(), mVox,
			                   kLabelInf, (const SolveControl *)ctl, kGateRelabelNew);
		}
		hipLaunchKernelGGL(k_fill_u8, dim3(blocks_for(mTiles, 256)), dim3(256), 0, mStream, mDirtyA.get(),
		                   (int64_t)mTiles, (uint8_t)0, (const SolveControl *)ctl, kGateRelabelNew);
		for (int c = 0; c < 2; ++c) {
			if (mSweep.size[c] > 0) {
				launch_bfs_pass(mSweep.tiles.get() + mSweep.start[c], &ctl->first[c], mSweep.size[c], 0,
				                mSweep.size[c] * tile_voxels, mConn == 6 ? c + 1 : 0);
			}
		}
		for (int cycle = 0; cycle < kGraphSweepCycles; ++cycle) {
			for (int c = 0; c < 2; ++c) {
				launch_compact(mSweep, mDirtyA.get(), c, true, nullptr, ctl, kGateRelabel);
				launch_bfs_pass(mList.get(), mListCount, mSweep.max_size, 0, -1, 0);
			}
		}
		hipLaunchKernelGGL(k_ctl_relabel_end, dim3(1), dim3(kCtlThreads), 0, mStream, ctl,
		                   (const uint8_t *)mDirtyA.get(), (const uint8_t *)mActive.get(), mTiles);
		// discharge phases, heaviest tiles first
		for (int p = 0; p < phases; ++p) {
			launch_compact(mPhase, mActive.get(), p % ncolours, false, mTileWork.get(), ctl, kGateDischarge);
			launch_discharge(max_iters, wide);
		}
		hipLaunchKernelGGL(k_ctl_discharge_end, dim3(1), dim3(64), 0, mStream, ctl, phases);
		// termination test (or the final S_min closure after a relabel found nothing active)
		hipLaunchKernelGGL(k_ctl_stage_begin, dim3(1), dim3(64), 0, mStream, ctl, kGateClosureNew, sweep_sizes);
		hipLaunchKernelGGL(k_closure_reset, dim3(blocks_for(mVox / 16 + 1, 256, 4096)), dim3(256), 0, mStream, mS.get(),
		                   mVox, (int32_t *)nullptr, mDirtyB.get(), (int64_t)mTiles, (const SolveControl *)ctl,
		                   kGateClosureNew);
		for (int c = 0; c < 2; ++c) {
			if (mSweep.size[c] > 0) {
				launch_closure_pass(mSweep.tiles.get() + mSweep.start[c], &ctl->first[c], mSweep.size[c], true,
				                    &ctl->reached, ctl, mSweep.size[c] * tile_voxels);
			}
		}
		for (int cycle = 0; cycle < kGraphSweepCycles; ++cycle) {
			for (int c = 0; c < 2; ++c) {
				launch_compact(mSweep, mDirtyB.get(), c, true, nullptr, ctl, kGateClosure | kGateUnlessReached);
				launch_closure_pass(mList.get(), mListCount, mSweep.max_size, false, &ctl->reached, ctl, -1);
			}
		}
		hipLaunchKernelGGL(k_ctl_closure_end, dim3(1), dim3(kCtlThreads), 0, mStream, ctl,
		                   (const uint8_t *)mDirtyB.get(), mTiles);
		// visit counters and per-grid flows, read back with the control block
		hipLaunchKernelGGL(k_sum_slots, dim3(blocks_for((int64_t)(3 + n), 256)), dim3(256), 0, mStream, mSlots.get(),
		                   (int64_t)(3 + n), totals);
		CGP_LAUNCH_CHECK();
		mCapturing = false;
		CGP_HIP(hipStreamEndCapture(mStream, &graph));
	} catch (...) {
		mCapturing = false;
		hipGraph_t partial = nullptr;
		(void)hipStreamEndCapture(mStream, &partial);
		if (partial) {
			(void)hipGraphDestroy(partial);
		}
		(void)hipGetLastError();
		throw;
	}
	hipGraphExec_t exec = nullptr;
	const hipError_t err = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
	(void)hipGraphDestroy(graph);
	CGP_HIP(err);
	mGraph.exec = exec;
	mGraph.generation = mGeneration;
	mGraph.phases = phases;
	mGraph.iterations = max_iters;
	mGraph.wide = wide;
}

int64_t GridSolver::solve_graph(int phases, int max_iters, int max_rounds, cgp_stats &stats)
{
	const int ncolours = mConn == 6 ? 2 : 8;
	// a replay starts its phases at colour 0, so every colour needs a phase in each replay
	phases = (phases + ncolours - 1) / ncolours * ncolours;
	const size_t n = mInst.size();
	bool wide = mWideResiduals;
	if (!mGraph.exec || mGraph.generation != mGeneration || mGraph.phases != phases ||
	    mGraph.iterations != max_iters || mGraph.wide != wide) {
		capture_round_graph(phases, max_iters, wide);
	}
	unsigned long long *pinned = reinterpret_cast<unsigned long long *>(mHostScratch.get());
	const SolveControl *host_ctl = reinterpret_cast<const SolveControl *>(pinned);
	int round_seen = -1, replays_in_round = 0;
	for (int replay = 0;; ++replay) {
		CGP_HIP(hipGraphLaunch(mGraph.exec, mStream));
		CGP_HIP(hipMemcpyAsync(pinned, mCtl.get(), (kCtlTotals + 3 + n) * sizeof(unsigned long long),
		                       hipMemcpyDeviceToHost, mStream));
		CGP_HIP(hipStreamSynchronize(mStream));
		++stats.graph_replays;
		const int mode = host_ctl->mode;
		if (mode == kModeFailed) {
			apply_status(host_ctl->status);  // throws
			fail(CGP_ERR_DEVICE, "device-driven solve failed with status " + std::to_string(host_ctl->status));
		}
		if (mode == kModeNeedWide) {
			// some edge's capacity pair exceeds 32767: stage the residuals as int32 from here on
			mWideResiduals = true;
			wide = true;
			capture_round_graph(phases, max_iters, true);
			hipLaunchKernelGGL(k_ctl_set_mode, dim3(1), dim3(64), 0, mStream, control(), (int32_t)kModeRelabelNew);
			CGP_LAUNCH_CHECK();
			continue;
		}
		if (mode == kModeNotConverged) {
			fail(CGP_ERR_NOT_CONVERGED, "max_rounds (" + std::to_string(max_rounds) + ") reached before convergence");
		}
		if (mode == kModeDoneClosed || mode == kModeDoneExact) {
			mClosedByCheck = mode == kModeDoneClosed;
			break;
		}
		// a global relabel or a closure takes a bounded number of colour-sweep passes; a round that
		// needs far more replays than any sweep can take has stopped making progress
		if (host_ctl->rounds != round_seen) {
			round_seen = host_ctl->rounds;
			replays_in_round = 0;
		}
		if (++replays_in_round > kMaxReplaysPerRound) {
			fail(CGP_ERR_NOT_CONVERGED, "device-driven solve made no progress: stage " + std::to_string(mode) +
			                                ", round " + std::to_string(host_ctl->rounds) + ", " +
			                                std::to_string(host_ctl->dirty) + " dirty tiles after " +
			                                std::to_string(replays_in_round) + " replays of the round");
		}
	}
	mWideResiduals = (host_ctl->status & 4) != 0;
	stats.rounds = host_ctl->rounds;
	stats.discharge_phases = host_ctl->phases;
	stats.discharge_tile_visits = (int64_t)pinned[kCtlTotals];
	stats.bfs_tile_visits = (int64_t)pinned[kCtlTotals + 1];
	stats.cut_tile_visits = (int64_t)pinned[kCtlTotals + 2];
	int64_t total = 0;
	for (size_t i = 0; i < n; ++i) {
		mFlows[i] = (int64_t)pinned[kCtlTotals + 3 + i];
		total += mFlows[i];
	}
	return total;
}

int64_t GridSolver::solve(const cgp_policy &policy, cgp_stats *stats_out, const TraceHook &trace)
{
	if (mVox == 0) {
		fail(CGP_ERR_STATE, "no grid set");
	}
	if (!std::isfinite(policy.minimal_residual) || policy.minimal_residual < 0.0f) {
		fail(CGP_ERR_INVALID, "minimal_residual must be finite and non-negative");
	}
	if (policy.engine != CGP_ENGINE_AUTO && policy.engine != CGP_ENGINE_TILES && policy.engine != CGP_ENGINE_GRAPH) {
		fail(CGP_ERR_INVALID, "unknown engine " + std::to_string(policy.engine));
	}
	cgp_stats stats;
	std::memset(&stats, 0, sizeof(stats));
	const auto t_start = std::chrono::steady_clock::now();
	mSolved = false;
	mStoppedEarly = false;
	mLabelsExact = false;
	mClosedByCheck = false;
	mStatusPending = false;
	if (mCertificate > 0 && mInst.size() != 1) {
		fail(CGP_ERR_INVALID, "a certificate applies to a single grid, not a batch");
	}
	for (KernelRecord &rec : mProf) {
		rec = KernelRecord();
	}
	mPending.clear();
	configure_tiles(pick_shape(policy));
	const int ncolours = mConn == 6 ? 2 : 8;
	const TunedPolicy tp = tuned_policy(mConn, mInst.size() > 1, mVox / (int64_t)mInst.size());
	const int phases = policy.discharges_per_relabel > 0 ? policy.discharges_per_relabel : tp.phases;
	const int max_iters = policy.max_tile_iterations > 0 ? policy.max_tile_iterations : tp.iterations;
	const int max_rounds = policy.max_rounds > 0 ? policy.max_rounds : 1 << 20;
	// (a threshold of 2^31 or more closes every edge, as INT32_MAX does)
	int32_t threshold = policy.minimal_residual > 0.0f
	                        ? (int32_t)std::min<double>(std::floor((double)policy.minimal_residual), 2147483647.0)
	                        : 0;

	// The device-driven engine takes no host decision per stage: it serves solves without trace
	// hooks, per-kernel profiling, an edge threshold or a certificate.
	// It is opt-in: the host-driven loop measured as fast or faster on every workload timed (256^3,
	// the 512-grid batch, 26-connected grids), so AUTO takes the host-driven loop. 26-connected
	// grids are refused: a 512x512x384 26-connected graph solve stopped answering twice, with no
	// cause found.
	const bool graph_ok = !trace.fn && !mProfile && threshold == 0 && mCertificate == 0;
	if (policy.engine == CGP_ENGINE_GRAPH && !graph_ok) {
		fail(CGP_ERR_INVALID, "the graph engine takes no trace hook, profiling, minimal_residual or certificate");
	}
	if (policy.engine == CGP_ENGINE_GRAPH && mConn != 6) {
		fail(CGP_ERR_INVALID, "the graph engine serves 6-connected grids; 26-connected grids run on the tiles engine");
	}
	const bool use_graph = policy.engine == CGP_ENGINE_GRAPH;
	// one launch clears the init status, the flow and visit-counter slots, the alternating counts of
	// the multi-workgroup compactions and (device-driven solve) the control block
	{
		const int64_t words = (int64_t)mSlots.size();
		hipLaunchKernelGGL(k_solve_reset, dim3(blocks_for(words, 256)), dim3(256), 0, mStream, mStatus.get(),
		                   mSlots.get(), words, mCursors.get(), kCursorWords, use_graph ? control() : nullptr,
		                   (int32_t)max_rounds);
		CGP_LAUNCH_CHECK();
		mOrderParity = mListParity = 0;
	}
	if (use_graph) {
		mStarted = true;
		mDirtyAClean = mDirtyBClean = false;  // the graph clears its flags itself
		const int64_t flow = solve_graph(phases, max_iters, max_rounds, stats);
		stats.flow = flow;
		stats.total_ms = ms_since(t_start);
		stats.voxels = mVox;
		stats.tile_x = mShape.tx;
		stats.tile_y = mShape.ty;
		stats.tile_z = mShape.tz;
		stats.conn = mConn;
		mSolved = true;
		mLabelsExact = !mClosedByCheck;
		if (stats_out) {
			*stats_out = stats;
		}
		return flow;
	}

	launch_init(nullptr);
	mStarted = true;
	if (trace.fn) {
		int32_t *pinned = reinterpret_cast<int32_t *>(mHostScratch.get());
		CGP_HIP(hipMemcpyAsync(pinned, mStatus.get(), sizeof(int32_t), hipMemcpyDeviceToHost, mStream));
		CGP_HIP(hipStreamSynchronize(mStream));
		apply_status(pinned[0]);
	} else {
		// read with the first count of the relabel (one host wait fewer); it comes before any
		// discharge, which needs the residual width
		mStatusPending = true;
	}
	stats.init_ms = ms_since(t_start);
	trace.emit(CGP_TRACE_STARTED, 0, 0, 0);

	int round = 0;
	int64_t phase = 0;  // colours rotate across rounds, so every colour is discharged whatever `phases` is
	// One-wave 6C tiles: when the last relabel ended at its first check, the next one queues its
	// discharge phases behind that check instead of waiting for it (one host wait fewer per round).
	// If the check then finds dirty tiles left, the phases ran on labels not yet exact; that only
	// costs work, since pushes along any labels keep a valid preflow and the closure test decides
	// termination; such a round goes to the closure test, and the next relabel waits again.
	const bool may_defer = mConn == 6 && mShape.tx * mShape.ty == 64 && !trace.fn && !mProfile && mCertificate == 0;
	bool exact_at_close = false;
	for (;;) {
		trace.emit(CGP_TRACE_BEGIN_ITERATION, round, 0, 0);
		auto t0 = std::chrono::steady_clock::now();
		const bool defer = may_defer && mRelabelOneCheck && threshold == 0 && round < max_rounds;
		mSweepChecks = 0;
		int active = run_global_relabel(threshold, stats, defer);
		if (defer) {
			// the phases and the termination closure follow in the same queue; the closure's first host
			// wait also brings the relabel's counts
			for (int p = 0; p < phases; ++p, ++phase) {
				run_discharge_phase((int)(phase % ncolours), max_iters);
			}
			stats.relabel_ms += ms_since(t0);
			++stats.rounds;
			stats.discharge_phases += phases;
			++round;
			t0 = std::chrono::steady_clock::now();
			const bool reached = run_source_closure(stats, true);
			stats.cut_ms += ms_since(t0);
			const int32_t *counts = deferred_counts();
			const int pending = counts[1];
			active = counts[2];
			mDirtyAClean = pending == 0;
			mRelabelOneCheck = pending == 0;
			learn_relabel_cycles(pending > 0 ? mRelabelCycles + 1 : 0);
			if (!reached) {
				// (with nothing active after an exact relabel, the labels are exact too)
				mClosedByCheck = true;
				exact_at_close = pending == 0 && active == 0;
				break;
			}
			continue;
		}
		mRelabelOneCheck = mSweepChecks == 1;
		stats.relabel_ms += ms_since(t0);
		++stats.rounds;
		trace.emit(CGP_TRACE_AFTER_RELABEL, round, 0, 0);
		if (active == 0) {
			if (threshold > 0) {
				threshold = 0;  // exact finish
				continue;
			}
			break;
		}
		if (round >= max_rounds) {
			fail(CGP_ERR_NOT_CONVERGED, "max_rounds (" + std::to_string(max_rounds) + ") reached before convergence");
		}
		t0 = std::chrono::steady_clock::now();
		bool certified = false;
		for (int p = 0; p < phases; ++p, ++phase) {
			run_discharge_phase((int)(phase % ncolours), max_iters);
			++stats.discharge_phases;
			if (mCertificate > 0 && read_flow() >= mCertificate) {
				// the flow into t reached the capacity of a cut the caller names: the preflow is
				// maximum if that is a true cut capacity, which the closure test below confirms
				certified = true;
				break;
			}
		}
		if (trace.fn) {
			trace.emit(CGP_TRACE_AFTER_PUSH, round, 0, read_flow());
		}
		stats.discharge_ms += ms_since(t0);
		++round;
		if (threshold == 0 || certified) {
			// termination test by the closure instead of one more global relabel: when the closure of
			// the excess reaches no voxel with a residual sink link, no excess can reach t, so the
			// preflow is maximum and the closure is S_min. Otherwise it stops at the first check that
			// sees t reached and the loop relabels as before (a certificate that was no cut capacity
			// is then dropped, and the solve runs to completion).
			t0 = std::chrono::steady_clock::now();
			const bool reached = run_source_closure(stats, true);
			stats.cut_ms += ms_since(t0);
			if (!reached) {
				mClosedByCheck = true;
				mStoppedEarly = certified;
				break;
			}
			if (certified) {
				mCertificate = 0;
			}
		}
	}
	auto t0 = std::chrono::steady_clock::now();
	if (!mClosedByCheck) {
		run_source_closure(stats);
	}
	unsigned long long visits[3];
	// the closure's last check normally brought the totals already
	const int64_t flow = mTotalsRead ? totals_from_readback(visits) : read_flow(visits);
	stats.cut_ms += ms_since(t0);
	stats.discharge_tile_visits = (int64_t)visits[0];
	stats.bfs_tile_visits = (int64_t)visits[1];
	stats.cut_tile_visits = (int64_t)visits[2];
	stats.flow = flow;
	stats.total_ms = ms_since(t_start);
	stats.voxels = mVox;
	stats.tile_x = mShape.tx;
	stats.tile_y = mShape.ty;
	stats.tile_z = mShape.tz;
	stats.conn = mConn;
	if (mProfile) {
		finish_profile(stats);
	}
	mSolved = true;
	// labels are those of the last relabel when the solve stopped at the closure test; T_min and
	// label readbacks relabel once on request
	mLabelsExact = !mClosedByCheck || exact_at_close;
	trace.emit(CGP_TRACE_AFTER_PUSH, round, 1, flow);
	trace.emit(CGP_TRACE_FINISHED, round, 1, flow);
	if (stats_out) {
		*stats_out = stats;
	}
	return flow;
}

// After a solve that stopped at a certificate the labels are those of an earlier relabel: one exact
// relabel of the final residual graph gives the vertices that still reach t.
void GridSolver::ensure_exact_labels()
{
	if (!mLabelsExact) {
		cgp_stats scratch;
		std::memset(&scratch, 0, sizeof(scratch));
		run_global_relabel(0, scratch);
		mLabelsExact = true;
	}
}

void GridSolver::sink_side(int index, uint8_t *host_out, uint8_t mask, uint8_t background)
{
	if (!mSolved) {
		fail(CGP_ERR_STATE, "fill_segments before max_flow");
	}
	ensure_exact_labels();
	const InstanceDesc &d = instance(index);
	DeviceBuffer<uint8_t> tmp(d.nvox);
	hipLaunchKernelGGL(k_sink_side_mask, dim3(blocks_for(d.nvox, 256, 8192)), dim3(256), 0, mStream,
	                   mH.get() + d.vbase, d.nvox, tmp.get(), mask, background);
	CGP_LAUNCH_CHECK();
	CGP_HIP(hipMemcpyAsync(host_out, tmp.get(), d.nvox, hipMemcpyDeviceToHost, mStream));
	CGP_HIP(hipStreamSynchronize(mStream));
}

void GridSolver::source_side(int index, uint8_t *host_out)
{
	if (!mSolved) {
		fail(CGP_ERR_STATE, "source_set before max_flow");
	}
	const InstanceDesc &d = instance(index);
	CGP_HIP(hipMemcpyAsync(host_out, mS.get() + d.vbase, d.nvox, hipMemcpyDeviceToHost, mStream));
	CGP_HIP(hipStreamSynchronize(mStream));
}

void GridSolver::source_side_device(int index, uint8_t *out)
{
	if (!mSolved) {
		fail(CGP_ERR_STATE, "source_set before max_flow");
	}
	const InstanceDesc &d = instance(index);
	CGP_HIP(hipMemcpyAsync(out, mS.get() + d.vbase, d.nvox, hipMemcpyDeviceToDevice, mStream));
	CGP_HIP(hipStreamSynchronize(mStream));
}

void GridSolver::labels(int index, int32_t *host_out)
{
	if (!mSolved) {
		fail(CGP_ERR_STATE, "labels before max_flow");
	}
	ensure_exact_labels();
	const InstanceDesc &d = instance(index);
	CGP_HIP(hipMemcpyAsync(host_out, mH.get() + d.vbase, d.nvox * sizeof(int32_t), hipMemcpyDeviceToHost, mStream));
	CGP_HIP(hipStreamSynchronize(mStream));
}

void GridSolver::state(int index, int32_t *excess, int32_t *labels, uint32_t *saturated)
{
	if (!mStarted) {
		fail(CGP_ERR_STATE, "no solve has started on this grid");
	}
	const InstanceDesc &d = instance(index);
	if (excess) {
		CGP_HIP(hipMemcpyAsync(excess, mE.get() + d.vbase, d.nvox * sizeof(int32_t), hipMemcpyDeviceToHost, mStream));
	}
	if (saturated) {
		DeviceBuffer<uint32_t> bits(d.nvox);
		const int half = mConn / 2;
		const int32_t *r = mR.get() + d.vbase * mConn;
		const int32_t *cn = mCn.get() + d.vbase * half;
		const int32_t *cnb = mHasBack ? mCnb.get() + d.vbase * half : nullptr;
		const int blocks = blocks_for(d.nvox, 256, 8192);
		if (mConn == 6) {
			hipLaunchKernelGGL(k_saturated_bits<6>, dim3(blocks), dim3(256), 0, mStream, d.nx, d.ny, d.nz, r, cn, cnb,
			                   bits.get());
		} else {
			hipLaunchKernelGGL(k_saturated_bits<26>, dim3(blocks), dim3(256), 0, mStream, d.nx, d.ny, d.nz, r, cn, cnb,
			                   bits.get());
		}
		CGP_LAUNCH_CHECK();
		CGP_HIP(hipMemcpyAsync(saturated, bits.get(), d.nvox * sizeof(uint32_t), hipMemcpyDeviceToHost, mStream));
		CGP_HIP(hipStreamSynchronize(mStream));
	}
	if (labels) {
		if (mSolved) {
			ensure_exact_labels();
		}
		CGP_HIP(hipMemcpyAsync(labels, mH.get() + d.vbase, d.nvox * sizeof(int32_t), hipMemcpyDeviceToHost, mStream));
	}
	CGP_HIP(hipStreamSynchronize(mStream));
}

void GridSolver::residuals(int index, int32_t *r, int32_t *rt)
{
	if (!mStarted) {
		fail(CGP_ERR_STATE, "no solve has started on this grid");
	}
	const InstanceDesc &d = instance(index);
	if (r) {
		CGP_HIP(hipMemcpyAsync(r, mR.get() + d.vbase * mConn, (size_t)d.nvox * mConn * sizeof(int32_t),
		                       hipMemcpyDeviceToHost, mStream));
	}
	if (rt) {
		CGP_HIP(hipMemcpyAsync(rt, mRt.get() + d.vbase, (size_t)d.nvox * sizeof(int32_t), hipMemcpyDeviceToHost,
		                       mStream));
	}
	CGP_HIP(hipStreamSynchronize(mStream));
}

GridSolver::DeviceGrid GridSolver::device_grid(int index) const
{
	if (mInst.empty()) {
		fail(CGP_ERR_STATE, "no grid set");
	}
	const InstanceDesc &d = instance(index);
	return DeviceGrid{d.nx, d.ny, d.nz, mCs.get() + d.vbase, mCt.get() + d.vbase};
}

const uint8_t *GridSolver::device_source_side(int index)
{
	if (!mSolved) {
		fail(CGP_ERR_STATE, "the grid is not solved");
	}
	const InstanceDesc &d = instance(index);
	return mS.get() + d.vbase;
}

// ---- no-alloc solve ------------------------------------------------------------------------------
// A grid solve that allocates nothing: every device buffer of the solver (per-voxel state, tile
// flags and lists, counters) is a piece of `scratch`, a file-scope device array of a fixed size
// that each GPU the library runs on holds until the process ends. In host code `scratch` names a
// host-side stand-in; the device address of each GPU's copy comes from hipGetSymbolAddress, called
// once per device with that device current, and only that address is ever used.

namespace {

__device__ __align__(256) uint8_t scratch[kNoAllocScratchBytes];

constexpr int kNoAllocMaxDevices = 64;

struct NoAllocDevice {
	std::mutex lock;           // calls on one device run one at a time
	bool ready = false;
	Arena arena;               // over this device's copy of `scratch`
	std::unique_ptr<GridSolver> solver;
};

NoAllocDevice g_noalloc[kNoAllocMaxDevices];

NoAllocDevice &noalloc_device(int device)
{
	if (device < 0 || device >= kNoAllocMaxDevices) {
		fail(CGP_ERR_DEVICE, "device " + std::to_string(device) + " out of range for the no-alloc entry point");
	}
	return g_noalloc[device];
}

}  // namespace

int64_t noalloc_max_flow(int nx, int ny, int nz, int conn, const int32_t *cs, const int32_t *ct, const int32_t *cn,
                         const int32_t *cnb, uint8_t *source_set, hipStream_t stream)
{
	int device = 0;
	CGP_HIP(hipGetDevice(&device));
	NoAllocDevice &nd = noalloc_device(device);
	std::lock_guard<std::mutex> guard(nd.lock);
	if (!nd.ready) {
		// set-up, once per device, with that device current: the address of its copy of the array
		void *ptr = nullptr;
		CGP_HIP(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
		nd.arena.base = static_cast<char *>(ptr);
		nd.arena.size = kNoAllocScratchBytes;
		nd.arena.used = 0;
		nd.solver.reset(new GridSolver(stream, &nd.arena));
		nd.ready = true;
	}
	GridSolver &solver = *nd.solver;
	solver.set_stream(stream);
	try {
		solver.set_grid_in_place(nx, ny, nz, conn, cs, ct, cn, cnb);
		cgp_policy policy = default_policy();
		policy.engine = CGP_ENGINE_TILES;  // plain launches: capturing a HIP graph may allocate
		cgp_stats stats;
		const int64_t flow = solver.solve(policy, &stats, TraceHook());
		if (source_set) {
			solver.source_side_device(0, source_set);
		}
		return flow;
	} catch (const Error &e) {
		if (e.code == CGP_ERR_INVALID && std::string(e.what()).rfind("scratch exhausted", 0) == 0) {
			fail(CGP_ERR_INVALID, "a " + std::to_string(nx) + "x" + std::to_string(ny) + "x" + std::to_string(nz) +
			                          " " + std::to_string(conn) + "-connected grid does not fit the " +
			                          std::to_string(kNoAllocScratchBytes >> 20) + " MiB scratch of the no-alloc "
			                          "entry point (" + e.what() + ")");
		}
		throw;
	}
}

}  // namespace cgp
