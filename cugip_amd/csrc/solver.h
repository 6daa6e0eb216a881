// Host-side solver classes behind the C ABI.
#pragma once

#include <cstdint>
#include <utility>
#include <vector>

#include "common.h"
#include "grid_control.h"
#include "grid_tile.h"

namespace cgp {

struct TraceHook {
	cgp_trace_fn fn = nullptr;
	void *user = nullptr;
	void emit(int event, int iteration, int done, int64_t flow, const int32_t *levels = nullptr, int count = 0) const
	{
		if (fn) {
			fn(user, event, iteration, done, flow, levels, count);
		}
	}
};

cgp_policy default_policy();
void default_volume_weights(cgp_volume_weights &w);

// One grid of a batch, as the caller passes it (layout of cgp_set_grid).
struct GridSpec {
	int nx, ny, nz;
	const int32_t *cs, *ct, *cn, *cnb;
};

// 3D grid graphs (6- or 26-connected), implicit edges, SoA int32 state in HBM. Holds one grid or a
// batch of independent grids of one connectivity, solved together: their tiles share the tile
// lists of every phase, so one launch serves all instances.
class GridSolver {
public:
	explicit GridSolver(hipStream_t stream) : mStream(stream) {}
	// A solver whose device memory all comes from `arena` (a fixed block it never frees), with
	// pageable host staging: it calls no allocating HIP routine. The memory of one grid is carved
	// afresh by every set_grid_in_place.
	GridSolver(hipStream_t stream, Arena *arena) : mStream(stream), mArena(arena) {}
	void set_stream(hipStream_t stream) { mStream = stream; }
	// A single grid whose capacities are device arrays the solver reads in place (no copies); they
	// must stay valid and unchanged while the solver uses them.
	void set_grid_in_place(int nx, int ny, int nz, int conn, const int32_t *cs, const int32_t *ct, const int32_t *cn,
	                       const int32_t *cnb);
	// The source side S_min of grid `index` after a solve, as device bytes (copied on the solver's
	// stream into `out`, device memory of its voxel count).
	void source_side_device(int index, uint8_t *out);

	void set_grid(int nx, int ny, int nz, int conn, const int32_t *cs, const int32_t *ct, const int32_t *cn,
	              const int32_t *cnb, bool device_pointers);
	void set_batch(int conn, const GridSpec *specs, int count, bool device_pointers);
	// Capacities computed on the device from a volume (and optional markers).
	void set_grid_volume(int nx, int ny, int nz, int conn, const uint8_t *volume, const uint8_t *markers,
	                     const cgp_volume_weights &weights, bool device_pointers);
	void capacities(int index, int32_t *cs, int32_t *ct, int32_t *cn, int32_t *cnb);
	void state(int index, int32_t *excess, int32_t *labels, uint32_t *saturated);
	void residuals(int index, int32_t *r, int32_t *rt);
	int64_t cut_capacity(int index, const uint8_t *host_side);
	int64_t cut_capacity_device(int index, const uint8_t *device_side);
	// Device views for the learned cut proposal.
	struct DeviceGrid {
		int nx, ny, nz;
		const int32_t *cs, *ct;
	};
	DeviceGrid device_grid(int index) const;
	const uint8_t *device_source_side(int index);
	// The volume of a grid set by set_grid_volume (kept on the device), else null.
	const uint8_t *device_volume() const { return mVolumeKept ? mVolume.get() : nullptr; }
	// A certificate > 0 (the capacity of some cut) lets the next solve stop as soon as the flow into
	// t reaches it; 0 disables it. Single grids only.
	void set_certificate(int64_t certificate) { mCertificate = certificate; }
	bool stopped_early() const { return mStoppedEarly; }
	int conn() const { return mConn; }
	const InstanceDesc &grid(int index) const { return instance(index); }
	// Returns the total flow of all instances; per-instance values in flows().
	int64_t solve(const cgp_policy &policy, cgp_stats *stats, const TraceHook &trace);
	const std::vector<int64_t> &flows() const { return mFlows; }
	void sink_side(int index, uint8_t *host_out, uint8_t mask, uint8_t background);
	void source_side(int index, uint8_t *host_out);
	void labels(int index, int32_t *host_out);
	int instance_count() const { return (int)mInst.size(); }
	int64_t instance_voxels(int index) const { return mInst[index].nvox; }
	int64_t vertex_count() const { return mVox; }
	bool solved() const { return mSolved; }
	void set_profiling(bool enable) { mProfile = enable; }
	int profile(cgp_kernel_profile *out, int capacity) const;
	~GridSolver();

private:
	enum Kern { kInit = 0, kBfs, kDischarge, kClosure, kCompact, kFill, kKernCount };
	struct KernelRecord {
		int launches = 0;
		int64_t voxels = 0;
		double ms = 0.0;
		double bytes_per_voxel = 0.0;
	};
	// Launch through here: with profiling on, the launch is bracketed by events and waited for.
	// voxels < 0 reads the voxel visits from the discharge tile counter.
	template <typename F>
	void launch(Kern kind, int64_t voxels, double bytes_per_voxel, F &&fn);

	struct Shape {
		int tx, ty, tz;
	};
	Shape pick_shape(const cgp_policy &policy);
	void configure_tiles(const Shape &shape);
	void launch_init(const SolveControl *ctl);
	int run_global_relabel(int32_t threshold, cgp_stats &stats, bool defer = false);
	// A colouring of the tiles: per-tile colour, the tile ids grouped by colour (then one count per
	// colour), the groups' starts and sizes, and the largest group.
	struct ColourSet {
		DeviceBuffer<uint8_t> colour;  // padded like the flag arrays
		DeviceBuffer<int32_t> tiles;
		std::vector<int> start, size;
		int colours = 0, max_size = 0;
	};
	void launch_compact(const ColourSet &set, uint8_t *flags, int colour, bool clear = false,
	                    const uint8_t *work = nullptr, const SolveControl *ctl = nullptr, unsigned gate = 0);
	void launch_bfs_pass(const int32_t *list, const int32_t *count, int grid, int32_t threshold, int64_t voxels,
	                     int first_mode);
	void launch_closure_pass(const int32_t *list, const int32_t *count, int grid, bool first, int32_t *reached,
	                         const SolveControl *ctl, int64_t voxels);
	void launch_discharge(int max_iters, bool wide);
	void count_two(uint8_t *a, uint8_t *b, int &na, int &nb, bool defer = false);
	void count_two_wait(int &na, int &nb, bool has_b);
	const int32_t *deferred_counts() const;
	bool mCheckDeferred = false;      // a count_two was queued without its host wait
	bool mRelabelOneCheck = true;     // the last relabel ended at its first check (the deferral pays)
	int mSweepChecks = 0;             // host checks of the last colour sweeps
	int mSweepCycles = 0;             // colour cycles (after the first sweep) of the last colour sweeps
	// cycles of a relabel's colour sweeps before their first check: what the last relabels needed
	// (tiles two rows high carry labels across y in twice the passes of four-row tiles), one fewer
	// after 8 relabels in a row ended at their first check; back to 2 for a new grid or tile shape
	int mRelabelCycles = 2, mRelabelFits = 0;
	static constexpr int kMaxRelabelCycles = 6;
	void learn_relabel_cycles(int needed);
	template <typename Pass>
	int colour_sweeps(uint8_t *dirty, uint8_t *also_count, int &also, Pass &&pass, bool *stop = nullptr,
	                  bool defer = false, int first_cycles = 2);
	void run_discharge_phase(int colour, int max_iters);
	// check: also report whether the closure reached a voxel with a residual sink link (the
	// preflow is then not maximum and the closure is not S_min); it stops early when it does
	bool run_source_closure(cgp_stats &stats, bool check = false);
	int64_t read_flow(unsigned long long *visits = nullptr);
	// Device-driven solve: init and one round of gated stages captured in a HIP graph, replayed with
	// one read of the control block per replay until it says done.
	struct RoundGraph {
		hipGraphExec_t exec = nullptr;
		int64_t generation = -1;
		int phases = 0, iterations = 0;
		bool wide = false;
		RoundGraph() = default;
		RoundGraph(const RoundGraph &) = delete;
		RoundGraph &operator=(const RoundGraph &) = delete;
		~RoundGraph();
		void reset();
	};
	static constexpr int kGraphSweepCycles = 3;  // colour-sweep cycles per stage and replay
	static constexpr int kMaxReplaysPerRound = 1024;
	static constexpr size_t kCtlTotals = sizeof(SolveControl) / sizeof(unsigned long long);
	void capture_round_graph(int phases, int max_iters, bool wide);
	int64_t solve_graph(int phases, int max_iters, int max_rounds, cgp_stats &stats);
	SolveControl *control() const { return reinterpret_cast<SolveControl *>(mCtl.get()); }
	const InstanceDesc &instance(int index) const;
	std::vector<InstanceDesc> describe(int conn, const GridSpec *specs, int count);
	// borrow (single grid): read the caller's device capacity arrays in place instead of copies
	void allocate_grids(int conn, const std::vector<InstanceDesc> &inst, bool has_back,
	                    const GridSpec *borrow = nullptr);

	hipStream_t mStream;
	Arena *mArena = nullptr;  // no-alloc mode: every device buffer is a piece of this block
	int mNx = 0, mNy = 0, mNz = 0, mConn = 0;  // dimensions of a single grid (instance 0)
	int64_t mVox = 0;                          // voxels of all instances
	std::vector<InstanceDesc> mInst;           // host copy of the instance table
	std::vector<int64_t> mFlows;               // per-instance flow of the last solve
	DeviceBuffer<InstanceDesc> mInstDev;       // batches only
	DeviceBuffer<int32_t> mTileInst;           // batches only: instance of every tile
	DeviceBuffer<int32_t> mInitBlockInst;      // batches only: instance of every init workgroup
	DeviceBuffer<int32_t> mInitFirstBlock;     // batches only: first init workgroup of every instance
	int mInitBlocks = 0;
	// configure_tiles' tables are kept while the grids and the tile shape stay the same
	bool mTilesReady = false;
	Shape mReadyShape{0, 0, 0};

	Shape mShape{0, 0, 0};
	// Discharge phases need that no two tiles of a colour touch (2 colours for 6C, 8 for 26C); the
	// relabel and closure sweeps only that a pass sees its face neighbours' results (face parity).
	ColourSet mPhase, mSweep;
	void build_colours(ColourSet &set, int ncolours, bool corners);
	int mNtx = 0, mNty = 0, mNtz = 0, mTiles = 0;
	bool mSolved = false;
	bool mStarted = false;  // a solve has initialised the residual state
	int64_t mCertificate = 0;
	DeviceBuffer<uint8_t> mVolume;  // volume of set_grid_volume
	bool mVolumeKept = false;
	bool mStoppedEarly = false;
	bool mLabelsExact = false;  // labels (T_min) come from an exact relabel of the final state
	bool mClosedByCheck = false;  // the last solve ended at the closure test (its S is S_min)
	bool mStatusPending = false;  // the init status is still to be read (with the next count_two)
	void apply_status(int32_t status);
	void ensure_exact_labels();
	bool mHasBack = false;
	bool mWideResiduals = true;  // some edge's capacity pair exceeds 32767: stage residuals as int32

	DeviceBuffer<int32_t> mCs, mCt, mCn, mCnb;
	DeviceBuffer<int32_t> mE, mRt, mH, mR;
	DeviceBuffer<uint8_t> mS;
	DeviceBuffer<uint32_t> mOmask;
	// slotted (kSlots per counter kind / per instance): views of mSlots, the 3 counters then the flows
	DeviceBuffer<unsigned long long> mSlots, mFlow, mCounters;
	DeviceBuffer<unsigned long long> mFlowTotal, mCounterTotal;
	DeviceBuffer<unsigned long long> mCutSlots;  // cut capacity: kSlots slots and their sum
	DeviceBuffer<int32_t> mStatus, mCount, mList, mAllTiles;
	DeviceBuffer<uint8_t> mActive, mDirtyA, mDirtyB;
	DeviceBuffer<uint8_t> mTileWork;  // per tile: discharge work estimate from the last relabel
	// multi-workgroup compactions: two sets of class cursors of the ordered lists, then two list
	// counts; each compaction uses one and clears the other (zeroed at the start of a solve)
	DeviceBuffer<int32_t> mCursors;
	static constexpr int kCountAcc = 2 * kWorkClasses + 2;  // then 3 words of k_count_flags2's accumulators
	static constexpr int kCursorWords = kCountAcc + 3;
	bool mDirtyAClean = false, mDirtyBClean = false;  // every relabel / closure tile flag is known to be 0
	int mOrderParity = 0, mListParity = 0;
	int mClassStride = 0;                   // class region size of the ordered lists
	const int32_t *mListCount = nullptr;    // count (or class cursors) of the last compacted list
	int mDischargeStride = 0;               // 0: the last list is contiguous; else its class stride
	DeviceBuffer<unsigned long long> mReadback;  // counts, visit-counter and flow totals (one read)
	// 2 words of counts (int32 [0..3]), the init status (int32 [4]), 3 words of visit counters
	static constexpr int kReadbackHead = 6;
	bool mPrefetchTotals = false;  // closure checks also read the flow and visit-counter totals
	bool mTotalsRead = false;      // the last check did (and nothing ran since)
	int64_t totals_from_readback(unsigned long long *visits);
	// device-driven solve: control block (kCtlTotals words), then 3 visit-counter totals and one flow
	// total per grid
	DeviceBuffer<unsigned long long> mCtl;
	RoundGraph mGraph;
	int64_t mGeneration = 0;  // bumped whenever buffers or tile tables a captured graph uses change
	bool mCapturing = false;
	PinnedBuffer<int64_t> mHostScratch;
	bool mProfile = false;
	KernelRecord mProf[kKernCount];
	// profiling: an event pair around every launch, read once the solve has ended (no host wait
	// between launches, so the timings are those of the unprofiled solve's kernels)
	struct PendingTiming {
		Kern kind;
		int64_t voxels;  // < 0: from the visit counters at the end
		double bytes_per_voxel;
		size_t event;
	};
	std::vector<hipEvent_t> mEvents;
	std::vector<PendingTiming> mPending;
	void finish_profile(const cgp_stats &stats);
};

// Scratch of the no-alloc grid solve, per device (CGP_NOALLOC_SCRATCH_BYTES in the C ABI).
constexpr size_t kNoAllocScratchBytes = (size_t)CGP_NOALLOC_SCRATCH_MIB << 20;

// Solves one grid whose capacities are device arrays, with the solver's device memory taken from
// a fixed per-device scratch array instead of allocations; writes S_min (device bytes, may be
// null) and returns the flow. Calls on one device are serialised.
int64_t noalloc_max_flow(int nx, int ny, int nz, int conn, const int32_t *cs, const int32_t *ct, const int32_t *cn,
                         const int32_t *cnb, uint8_t *source_set, hipStream_t stream);

// Region adjacency graph of the reference's region benchmarks (region_graph.cpp).
struct RegionGraph {
	int32_t vertices = 0;
	std::vector<int32_t> a, b;  // 0-based region ids of each edge record
	std::vector<double> sum;    // its boundary statistic
};
RegionGraph read_region_graph(const char *path);
// (object marker?, 0-based region id) per marker line
std::vector<std::pair<bool, int32_t>> read_region_markers(const char *path);
void region_graph_arrays(const RegionGraph &g, const std::vector<std::pair<bool, int32_t>> &markers,
                         double weight_scale, int32_t marker_capacity, cgp_edge *edges, int32_t *weights,
                         int32_t *cap_source, int32_t *cap_sink);

// Generic graph (CSR), the data layout of the reference's Graph<TFlow>.
class CsrSolver {
public:
	explicit CsrSolver(hipStream_t stream) : mStream(stream) {}

	void set_vertex_count(int n);
	void set_nweights(int m, const cgp_edge *edges, const int32_t *fwd, const int32_t *bwd);
	void set_tweights(const int32_t *source, const int32_t *sink);
	int64_t solve(const cgp_policy &policy, cgp_stats *stats, const TraceHook &trace);
	void sink_side(uint8_t *host_out, uint8_t mask, uint8_t background);
	void source_side(uint8_t *host_out);
	void labels(int32_t *host_out);
	int64_t vertex_count() const { return mN; }
	bool solved() const { return mSolved; }
	bool ready() const { return mN > 0 && mHaveTweights; }
	// Directed edges with positive capacity (tail, head, capacity) and the t-links, for export.
	void export_graph(std::vector<int32_t> &tail, std::vector<int32_t> &head, std::vector<int32_t> &cap,
	                  std::vector<int32_t> &cs, std::vector<int32_t> &ct);

private:
	int bfs(int32_t threshold, std::vector<int32_t> &level_starts, cgp_stats &stats);
	bool push_sweep(const std::vector<int32_t> &level_starts, cgp_stats &stats);
	void source_closure(cgp_stats &stats);
	int32_t read_i32(const int32_t *device_ptr);
	int64_t read_flow();

	hipStream_t mStream;
	int mN = 0;
	int64_t mArcs = 0;
	bool mSolved = false;
	bool mHaveTweights = false;
	std::vector<int32_t> mHostFirst;   // n + 1
	std::vector<int64_t> mIncoming;    // sum of capacities into each vertex
	std::vector<int32_t> mHostSource;  // source t-links (headroom check)
	DeviceBuffer<int32_t> mFirst, mHead, mMate, mArcCap, mRes;
	DeviceBuffer<int32_t> mCs, mCt, mE, mRt, mH, mQueue;
	DeviceBuffer<int32_t> mS;          // S_min marks
	DeviceBuffer<int32_t> mTail, mPushed, mStatus, mLevelStarts;
	DeviceBuffer<unsigned long long> mFlow;
	PinnedBuffer<int32_t> mPinned;
};

}  // namespace cgp
