// Generic-graph push-relabel on a CSR adjacency: the data layout and the
// algorithm of the reference's Graph<TFlow>
// (include/cugip/advanced_operations/graph_cut.hpp:213-302, detail/graph_cut_*.hpp),
// restated for exact int32/int64 arithmetic and finite sink capacities.
//
//   relabel: level-synchronous BFS from the sink frontier (level 1 = residual sink link)
//            over reverse residual edges, vertices appended to one queue in level order
//            (initBFSKernel, propagateFromVertex, bfs_iteration / bfs_multi_iteration,
//            graph_cut_relabeling.hpp:47-112, 1004-1240);
//   push:    levels from the top down to 1, every vertex of the level pushing to any neighbour
//            with a lower label, level 1 also into t (pushImplementation / PushIteration,
//            graph_cut_push.hpp:160-186, 395-482);
//   loop:    until a sweep pushes nothing (MinCut::compute, graph_cut_implementation.hpp:122-198).
// Small levels run inside one workgroup for many levels per launch, like the reference's
// single-block multi-level kernels.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>

#include "common.h"
#include "solver.h"

namespace cgp {

namespace {

constexpr int kMultiLevelLimit = 4096;   // level size handled by the single-workgroup kernels
constexpr int kMultiLevelThreads = 1024;
constexpr int kMaxLevelsPerLaunch = 1024;

__global__ __launch_bounds__(256) void k_csr_init_vertices(int n, const int32_t *cs, const int32_t *ct, int32_t *e,
                                                           int32_t *rt)
{
	for (int v = blockIdx.x * blockDim.x + threadIdx.x; v < n; v += gridDim.x * blockDim.x) {
		e[v] = cs[v];
		rt[v] = ct[v];
	}
}

__global__ __launch_bounds__(256) void k_csr_copy(int64_t n, const int32_t *src, int32_t *dst)
{
	for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
		dst[i] = src[i];
	}
}

// Append every vertex satisfying `pred` to queue (one atomic per wave).
__device__ inline void wave_append(bool pred, int v, int32_t *queue, int32_t *tail)
{
	const unsigned long long ballot = __ballot(pred);
	const int lane = threadIdx.x & 63;
	int base = 0;
	if (lane == 0 && ballot) {
		base = atomicAdd(tail, (int)__popcll(ballot));
	}
	base = __shfl(base, 0, 64);
	if (pred) {
		queue[base + (int)__popcll(ballot & ((1ull << lane) - 1ull))] = v;
	}
}

// Level 1 of the BFS: vertices whose sink link still has residual capacity.
__global__ __launch_bounds__(256) void k_csr_bfs_seed(int n, const int32_t *rt, int32_t threshold, int32_t *h,
                                                      int32_t *queue, int32_t *tail)
{
	const int stride = gridDim.x * blockDim.x;
	const int start = blockIdx.x * blockDim.x + threadIdx.x;
	// whole waves iterate together so that the ballot sees every lane
	const int rounds = (n + stride - 1) / stride;
	for (int r = 0; r < rounds; ++r) {
		const int v = start + r * stride;
		bool seed = false;
		if (v < n) {
			seed = rt[v] > threshold;
			h[v] = seed ? 1 : kLabelInf;
		}
		wave_append(seed, v, queue, tail);
	}
}

// Append v to queue where pred holds, among the lanes active at the call (also inside loops whose
// trip counts differ per lane): the lowest active lane reserves the entries of the whole wave with
// one atomic, and each lane stores at its rank.
__device__ inline void wave_append_active(bool pred, int v, int32_t *queue, int32_t *tail)
{
	const unsigned long long active = __ballot(1);
	const unsigned long long ballot = __ballot(pred);
	if (!ballot) {
		return;  // (the same for every active lane)
	}
	const int lane = threadIdx.x & 63;
	const int leader = __ffsll((long long)active) - 1;
	int base = 0;
	if (lane == leader) {
		base = atomicAdd(tail, (int)__popcll(ballot));
	}
	base = __shfl(base, leader, 64);
	if (pred) {
		queue[base + (int)__popcll(ballot & ((1ull << lane) - 1ull))] = v;
	}
}

// Label the unlabelled neighbours u of frontier vertex v that can reach v (residual u -> v > threshold).
__device__ inline void expand_towards_sink(int v, int32_t level, const int32_t *first, const int32_t *head,
                                           const int32_t *mate, const int32_t *res, int32_t threshold, int32_t *h,
                                           int32_t *queue, int32_t *tail)
{
	for (int a = first[v]; a < first[v + 1]; ++a) {
		const int u = head[a];
		bool claimed = false;
		if (h[u] == kLabelInf && res[mate[a]] > threshold) {
			claimed = atomicCAS(&h[u], kLabelInf, level + 1) == kLabelInf;
		}
		wave_append_active(claimed, u, queue, tail);
	}
}

// The flow into t is summed in kSlots slots on separate cache lines, picked by workgroup: one
// word taking an atomic from every workgroup serialises them in the L2.
constexpr int kFlowSlots = 32;
constexpr int kFlowSlotStride = 16;

__device__ inline void add_flow(unsigned long long *flow, long long total)
{
	if (total) {
		atomicAdd(flow + (blockIdx.x & (kFlowSlots - 1)) * kFlowSlotStride, (unsigned long long)total);
	}
}

__global__ void k_csr_sum_flow(const unsigned long long *slots, unsigned long long *total)
{
	if (threadIdx.x == 0 && blockIdx.x == 0) {
		unsigned long long sum = 0;
		for (int j = 0; j < kFlowSlots; ++j) {
			sum += slots[j * kFlowSlotStride];
		}
		*total = sum;
	}
}

__global__ __launch_bounds__(256) void k_csr_bfs_level(int32_t qs, int32_t qe, int32_t level, const int32_t *first,
                                                       const int32_t *head, const int32_t *mate, const int32_t *res,
                                                       int32_t threshold, int32_t *h, int32_t *queue, int32_t *tail)
{
	for (int i = qs + blockIdx.x * blockDim.x + threadIdx.x; i < qe; i += gridDim.x * blockDim.x) {
		expand_towards_sink(queue[i], level, first, head, mate, res, threshold, h, queue, tail);
	}
}

// Many small BFS levels in one workgroup. Writes the queue end after each level to starts[] and
// the number of levels processed to *levels_out.
__global__ __launch_bounds__(kMultiLevelThreads) void k_csr_bfs_multi(int32_t qs, int32_t qe, int32_t level,
                                                                      const int32_t *first, const int32_t *head,
                                                                      const int32_t *mate, const int32_t *res,
                                                                      int32_t threshold, int32_t *h, int32_t *queue,
                                                                      int32_t *tail, int32_t *starts,
                                                                      int32_t *levels_out)
{
	int done = 0;
	for (;;) {
		for (int i = qs + (int)threadIdx.x; i < qe; i += blockDim.x) {
			expand_towards_sink(queue[i], level, first, head, mate, res, threshold, h, queue, tail);
		}
		__syncthreads();
		const int32_t nt = atomicAdd(tail, 0);
		if (threadIdx.x == 0) {
			starts[done] = nt;
		}
		++done;
		__syncthreads();
		if (nt == qe || nt - qe > kMultiLevelLimit || done >= kMaxLevelsPerLaunch) {
			break;
		}
		qs = qe;
		qe = nt;
		++level;
	}
	if (threadIdx.x == 0) {
		*levels_out = done;
	}
}

__device__ inline bool push_vertex(int v, int32_t level, const int32_t *first, const int32_t *head,
                                   const int32_t *mate, int32_t *res, int32_t *e, int32_t *rt, const int32_t *h,
                                   long long &sink)
{
	int32_t ex = e[v];
	if (ex <= 0) {
		return false;
	}
	bool pushed = false;
	if (rt[v] > 0) {
		const int32_t f = ex < rt[v] ? ex : rt[v];
		rt[v] -= f;
		ex -= f;
		sink += f;
		pushed = true;
	}
	for (int a = first[v]; a < first[v + 1] && ex > 0; ++a) {
		const int u = head[a];
		const int32_t hu = h[u];
		const int32_t rv = res[a];
		if (hu < level && rv > 0) {
			const int32_t f = ex < rv ? ex : rv;
			res[a] = rv - f;
			res[mate[a]] += f;   // only v pushes along this edge during this level
			atomicAdd(&e[u], f); // several vertices of the level may feed u
			ex -= f;
			pushed = true;
		}
	}
	e[v] = ex;
	return pushed;
}

__global__ __launch_bounds__(256) void k_csr_push_level(int32_t qs, int32_t qe, int32_t level, const int32_t *queue,
                                                        const int32_t *first, const int32_t *head,
                                                        const int32_t *mate, int32_t *res, int32_t *e, int32_t *rt,
                                                        const int32_t *h, unsigned long long *flow, int32_t *pushed)
{
	__shared__ long long red[4];
	long long sink = 0;
	bool any = false;
	for (int i = qs + blockIdx.x * blockDim.x + threadIdx.x; i < qe; i += gridDim.x * blockDim.x) {
		any |= push_vertex(queue[i], level, first, head, mate, res, e, rt, h, sink);
	}
	if (any) {
		*pushed = 1;
	}
	const long long total = block_sum_i64<256>(sink, red);
	if (threadIdx.x == 0) {
		add_flow(flow, total);
	}
}

// Consecutive small levels, from `level` downwards, in one workgroup; stops at level 0 or at the
// first level above the size limit. Writes the first level not processed to *next_level.
__global__ __launch_bounds__(kMultiLevelThreads) void k_csr_push_multi(int32_t level, const int32_t *starts,
                                                                       const int32_t *queue, const int32_t *first,
                                                                       const int32_t *head, const int32_t *mate,
                                                                       int32_t *res, int32_t *e, int32_t *rt,
                                                                       const int32_t *h, unsigned long long *flow,
                                                                       int32_t *pushed, int32_t *next_level)
{
	__shared__ long long red[kMultiLevelThreads / 64];
	long long sink = 0;
	bool any = false;
	while (level >= 1) {
		const int32_t qs = starts[level - 1], qe = starts[level];
		if (qe - qs > kMultiLevelLimit) {
			break;
		}
		for (int i = qs + (int)threadIdx.x; i < qe; i += blockDim.x) {
			any |= push_vertex(queue[i], level, first, head, mate, res, e, rt, h, sink);
		}
		__syncthreads();
		--level;
	}
	if (any) {
		*pushed = 1;
	}
	const long long total = block_sum_i64<kMultiLevelThreads>(sink, red);
	if (threadIdx.x == 0) {
		add_flow(flow, total);
		*next_level = level;
	}
}

// S_min closure: mark vertices with excess, then everything reachable over residual arcs.
__global__ __launch_bounds__(256) void k_csr_closure_seed(int n, const int32_t *e, int32_t *mark, int32_t *queue,
                                                          int32_t *tail)
{
	const int stride = gridDim.x * blockDim.x;
	const int start = blockIdx.x * blockDim.x + threadIdx.x;
	const int rounds = (n + stride - 1) / stride;
	for (int r = 0; r < rounds; ++r) {
		const int v = start + r * stride;
		bool seed = false;
		if (v < n) {
			seed = e[v] > 0;
			mark[v] = seed ? 1 : 0;
		}
		wave_append(seed, v, queue, tail);
	}
}

__global__ __launch_bounds__(256) void k_csr_closure_level(int32_t qs, int32_t qe, const int32_t *first,
                                                           const int32_t *head, const int32_t *res, int32_t *mark,
                                                           int32_t *queue, int32_t *tail)
{
	for (int i = qs + blockIdx.x * blockDim.x + threadIdx.x; i < qe; i += gridDim.x * blockDim.x) {
		const int v = queue[i];
		for (int a = first[v]; a < first[v + 1]; ++a) {
			const int u = head[a];
			const bool claimed = res[a] > 0 && mark[u] == 0 && atomicCAS(&mark[u], 0, 1) == 0;
			wave_append_active(claimed, u, queue, tail);
		}
	}
}

__global__ __launch_bounds__(256) void k_csr_masks(int n, const int32_t *h, const int32_t *mark, uint8_t *out,
                                                   int mode, uint8_t mask, uint8_t background)
{
	for (int v = blockIdx.x * blockDim.x + threadIdx.x; v < n; v += gridDim.x * blockDim.x) {
		if (mode == 0) {
			out[v] = h[v] < kLabelInf ? mask : background;
		} else {
			out[v] = mark[v] ? 1 : 0;
		}
	}
}

int grid_blocks(int64_t n, int threads = 256, int cap = 8192)
{
	int64_t b = (n + threads - 1) / threads;
	return (int)std::max<int64_t>(1, std::min<int64_t>(b, cap));
}

double ms_since(std::chrono::steady_clock::time_point t)
{
	return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t).count();
}

}  // namespace

void CsrSolver::set_vertex_count(int n)
{
	if (n <= 0) {
		fail(CGP_ERR_INVALID, "vertex count must be positive");
	}
	mN = n;
	mArcs = 0;
	mSolved = false;
	mHaveTweights = false;
	mHostFirst.assign(n + 1, 0);
	mFirst.allocate(n + 1);
	CGP_HIP(hipMemsetAsync(mFirst.get(), 0, (n + 1) * sizeof(int32_t), mStream));
	mCs.allocate(n);
	mCt.allocate(n);
	mE.allocate(n);
	mRt.allocate(n);
	mH.allocate(n);
	mQueue.allocate(n);
	mS.allocate(n);
	mTail.allocate(1);
	mPushed.allocate(1);
	mStatus.allocate(1);
	mLevelStarts.allocate(kMaxLevelsPerLaunch + 2);
	mFlow.allocate(kFlowSlots * kFlowSlotStride + 1);  // slots, then their sum
	mPinned.allocate(kMaxLevelsPerLaunch + 8);
	mHead.release();
	mMate.release();
	mArcCap.release();
	mRes.release();
	CGP_HIP(hipStreamSynchronize(mStream));
}

void CsrSolver::set_nweights(int m, const cgp_edge *edges, const int32_t *fwd, const int32_t *bwd)
{
	if (mN <= 0) {
		fail(CGP_ERR_STATE, "set_nweights before set_vertex_count");
	}
	if (m < 0 || (m > 0 && (!edges || !fwd || !bwd))) {
		fail(CGP_ERR_INVALID, "invalid edge arrays");
	}
	// every edge makes two arcs, and arc indices are int32
	if (m >= (1 << 30)) {
		fail(CGP_ERR_INVALID, "at most 2^30 - 1 edges are supported");
	}
	std::vector<int32_t> degree(mN + 1, 0);
	for (int i = 0; i < m; ++i) {
		const uint32_t a = edges[i].first, b = edges[i].second;
		if (a >= (uint32_t)mN || b >= (uint32_t)mN) {
			fail(CGP_ERR_INVALID, "edge " + std::to_string(i) + " references a vertex outside [0, n)");
		}
		if (fwd[i] < 0 || bwd[i] < 0) {
			fail(CGP_ERR_INVALID, "edge " + std::to_string(i) + " has a negative capacity");
		}
		if ((int64_t)fwd[i] + bwd[i] > 0x7fffffffLL) {
			fail(CGP_ERR_OVERFLOW, "edge " + std::to_string(i) + ": fwd + bwd does not fit int32");
		}
		if (a == b) {
			continue;  // a loop carries no s-t flow
		}
		++degree[a];
		++degree[b];
	}
	std::vector<int32_t> first(mN + 1, 0);
	for (int v = 0; v < mN; ++v) {
		first[v + 1] = first[v] + degree[v];
	}
	const int64_t arcs = first[mN];
	std::vector<int32_t> pos(first.begin(), first.end() - 1);
	std::vector<int32_t> head(arcs), mate(arcs), cap(arcs);
	std::vector<int64_t> incoming(mN, 0);
	for (int i = 0; i < m; ++i) {
		const int a = (int)edges[i].first, b = (int)edges[i].second;
		if (a == b) {
			continue;
		}
		const int ia = pos[a]++, ib = pos[b]++;
		head[ia] = b;
		cap[ia] = fwd[i];
		mate[ia] = ib;
		head[ib] = a;
		cap[ib] = bwd[i];
		mate[ib] = ia;
		incoming[b] += fwd[i];
		incoming[a] += bwd[i];
	}
	mIncoming.swap(incoming);
	mHostFirst.swap(first);
	mArcs = arcs;
	mSolved = false;
	mFirst.allocate(mN + 1);
	CGP_HIP(hipMemcpyAsync(mFirst.get(), mHostFirst.data(), (mN + 1) * sizeof(int32_t), hipMemcpyHostToDevice, mStream));
	if (arcs > 0) {
		mHead.allocate(arcs);
		mMate.allocate(arcs);
		mArcCap.allocate(arcs);
		mRes.allocate(arcs);
		CGP_HIP(hipMemcpyAsync(mHead.get(), head.data(), arcs * sizeof(int32_t), hipMemcpyHostToDevice, mStream));
		CGP_HIP(hipMemcpyAsync(mMate.get(), mate.data(), arcs * sizeof(int32_t), hipMemcpyHostToDevice, mStream));
		CGP_HIP(hipMemcpyAsync(mArcCap.get(), cap.data(), arcs * sizeof(int32_t), hipMemcpyHostToDevice, mStream));
	}
	CGP_HIP(hipStreamSynchronize(mStream));
}

void CsrSolver::set_tweights(const int32_t *source, const int32_t *sink)
{
	if (mN <= 0) {
		fail(CGP_ERR_STATE, "set_tweights before set_vertex_count");
	}
	if (!source || !sink) {
		fail(CGP_ERR_INVALID, "t-link arrays must not be null");
	}
	for (int v = 0; v < mN; ++v) {
		if (source[v] < 0 || sink[v] < 0) {
			fail(CGP_ERR_INVALID, "vertex " + std::to_string(v) + " has a negative t-link capacity");
		}
	}
	mHostSource.assign(source, source + mN);
	CGP_HIP(hipMemcpyAsync(mCs.get(), source, mN * sizeof(int32_t), hipMemcpyHostToDevice, mStream));
	CGP_HIP(hipMemcpyAsync(mCt.get(), sink, mN * sizeof(int32_t), hipMemcpyHostToDevice, mStream));
	CGP_HIP(hipStreamSynchronize(mStream));
	mHaveTweights = true;
	mSolved = false;
}

void CsrSolver::export_graph(std::vector<int32_t> &tail, std::vector<int32_t> &head, std::vector<int32_t> &cap,
                             std::vector<int32_t> &cs, std::vector<int32_t> &ct)
{
	if (!ready()) {
		fail(CGP_ERR_STATE, "no complete graph set");
	}
	std::vector<int32_t> heads((size_t)mArcs), caps((size_t)mArcs);
	if (mArcs) {
		CGP_HIP(hipMemcpyAsync(heads.data(), mHead.get(), mArcs * sizeof(int32_t), hipMemcpyDeviceToHost, mStream));
		CGP_HIP(hipMemcpyAsync(caps.data(), mArcCap.get(), mArcs * sizeof(int32_t), hipMemcpyDeviceToHost, mStream));
	}
	cs.resize((size_t)mN);
	ct.resize((size_t)mN);
	CGP_HIP(hipMemcpyAsync(cs.data(), mCs.get(), mN * sizeof(int32_t), hipMemcpyDeviceToHost, mStream));
	CGP_HIP(hipMemcpyAsync(ct.data(), mCt.get(), mN * sizeof(int32_t), hipMemcpyDeviceToHost, mStream));
	CGP_HIP(hipStreamSynchronize(mStream));
	tail.clear();
	head.clear();
	cap.clear();
	for (int v = 0; v < mN; ++v) {
		for (int a = mHostFirst[(size_t)v]; a < mHostFirst[(size_t)v + 1]; ++a) {
			if (caps[(size_t)a] > 0) {
				tail.push_back(v);
				head.push_back(heads[(size_t)a]);
				cap.push_back(caps[(size_t)a]);
			}
		}
	}
}

int64_t CsrSolver::read_flow()
{
	unsigned long long *total = mFlow.get() + kFlowSlots * kFlowSlotStride;
	hipLaunchKernelGGL(k_csr_sum_flow, dim3(1), dim3(64), 0, mStream, (const unsigned long long *)mFlow.get(), total);
	CGP_LAUNCH_CHECK();
	CGP_HIP(hipMemcpyAsync(mPinned.get(), total, sizeof(unsigned long long), hipMemcpyDeviceToHost, mStream));
	CGP_HIP(hipStreamSynchronize(mStream));
	int64_t f;
	std::memcpy(&f, mPinned.get(), sizeof(f));
	return f;
}

int32_t CsrSolver::read_i32(const int32_t *device_ptr)
{
	CGP_HIP(hipMemcpyAsync(mPinned.get(), device_ptr, sizeof(int32_t), hipMemcpyDeviceToHost, mStream));
	CGP_HIP(hipStreamSynchronize(mStream));
	return mPinned[0];
}

int CsrSolver::bfs(int32_t threshold, std::vector<int32_t> &level_starts, cgp_stats &stats)
{
	CGP_HIP(hipMemsetAsync(mTail.get(), 0, sizeof(int32_t), mStream));
	hipLaunchKernelGGL(k_csr_bfs_seed, dim3(grid_blocks(mN)), dim3(256), 0, mStream, mN, mRt.get(), threshold, mH.get(),
	                   mQueue.get(), mTail.get());
	CGP_LAUNCH_CHECK();
	level_starts.assign(1, 0);
	level_starts.push_back(read_i32(mTail.get()));
	int level = 1;
	while (level_starts[level] > level_starts[level - 1]) {
		const int32_t qs = level_starts[level - 1], qe = level_starts[level];
		if (qe - qs <= kMultiLevelLimit) {
			hipLaunchKernelGGL(k_csr_bfs_multi, dim3(1), dim3(kMultiLevelThreads), 0, mStream, qs, qe, level, mFirst.get(),
			                   mHead.get(), mMate.get(), mRes.get(), threshold, mH.get(), mQueue.get(), mTail.get(),
			                   mLevelStarts.get(), mLevelStarts.get() + kMaxLevelsPerLaunch + 1);
			CGP_LAUNCH_CHECK();
			CGP_HIP(hipMemcpyAsync(mPinned.get(), mLevelStarts.get(), (kMaxLevelsPerLaunch + 2) * sizeof(int32_t),
			                       hipMemcpyDeviceToHost, mStream));
			CGP_HIP(hipStreamSynchronize(mStream));
			const int done = mPinned[kMaxLevelsPerLaunch + 1];
			stats.bfs_passes += done;
			for (int i = 0; i < done; ++i) {
				const int32_t nt = mPinned[i];
				if (nt == level_starts.back()) {
					return (int)level_starts.size() - 1;
				}
				level_starts.push_back(nt);
				++level;
			}
		} else {
			hipLaunchKernelGGL(k_csr_bfs_level, dim3(grid_blocks(qe - qs)), dim3(256), 0, mStream, qs, qe, level,
			                   mFirst.get(), mHead.get(), mMate.get(), mRes.get(), threshold, mH.get(), mQueue.get(),
			                   mTail.get());
			CGP_LAUNCH_CHECK();
			++stats.bfs_passes;
			const int32_t nt = read_i32(mTail.get());
			if (nt == qe) {
				break;
			}
			level_starts.push_back(nt);
			++level;
		}
	}
	return (int)level_starts.size() - 1;
}

bool CsrSolver::push_sweep(const std::vector<int32_t> &level_starts, cgp_stats &stats)
{
	const int top = (int)level_starts.size() - 1;
	DeviceBuffer<int32_t> starts(level_starts.size());
	CGP_HIP(hipMemcpyAsync(starts.get(), level_starts.data(), level_starts.size() * sizeof(int32_t),
	                       hipMemcpyHostToDevice, mStream));
	CGP_HIP(hipMemsetAsync(mPushed.get(), 0, sizeof(int32_t), mStream));
	int level = top;
	while (level >= 1) {
		const int32_t qs = level_starts[level - 1], qe = level_starts[level];
		if (qe - qs <= kMultiLevelLimit) {
			hipLaunchKernelGGL(k_csr_push_multi, dim3(1), dim3(kMultiLevelThreads), 0, mStream, level, starts.get(),
			                   mQueue.get(), mFirst.get(), mHead.get(), mMate.get(), mRes.get(), mE.get(), mRt.get(),
			                   mH.get(), mFlow.get(), mPushed.get(), mStatus.get());
			CGP_LAUNCH_CHECK();
			const int next = read_i32(mStatus.get());
			stats.discharge_phases += level - next;
			level = next;
		} else {
			hipLaunchKernelGGL(k_csr_push_level, dim3(grid_blocks(qe - qs)), dim3(256), 0, mStream, qs, qe, level,
			                   mQueue.get(), mFirst.get(), mHead.get(), mMate.get(), mRes.get(), mE.get(), mRt.get(),
			                   mH.get(), mFlow.get(), mPushed.get());
			CGP_LAUNCH_CHECK();
			++stats.discharge_phases;
			--level;
		}
	}
	return read_i32(mPushed.get()) != 0;
}

void CsrSolver::source_closure(cgp_stats &stats)
{
	CGP_HIP(hipMemsetAsync(mTail.get(), 0, sizeof(int32_t), mStream));
	hipLaunchKernelGGL(k_csr_closure_seed, dim3(grid_blocks(mN)), dim3(256), 0, mStream, mN, mE.get(), mS.get(),
	                   mQueue.get(), mTail.get());
	CGP_LAUNCH_CHECK();
	int32_t qs = 0, qe = read_i32(mTail.get());
	while (qe > qs) {
		hipLaunchKernelGGL(k_csr_closure_level, dim3(grid_blocks(qe - qs)), dim3(256), 0, mStream, qs, qe, mFirst.get(),
		                   mHead.get(), mRes.get(), mS.get(), mQueue.get(), mTail.get());
		CGP_LAUNCH_CHECK();
		++stats.cut_passes;
		qs = qe;
		qe = read_i32(mTail.get());
	}
}

int64_t CsrSolver::solve(const cgp_policy &policy, cgp_stats *stats_out, const TraceHook &trace)
{
	if (!ready()) {
		fail(CGP_ERR_STATE, "max_flow needs set_vertex_count and set_tweights first");
	}
	if (policy.minimal_residual < 0.0f || !std::isfinite(policy.minimal_residual)) {
		fail(CGP_ERR_INVALID, "minimal_residual must be finite and non-negative");
	}
	for (int v = 0; v < mN; ++v) {
		const int64_t in = (v < (int)mIncoming.size() ? mIncoming[v] : 0) + mHostSource[v];
		if (in > 0x7fffffffLL) {
			fail(CGP_ERR_OVERFLOW, "vertex " + std::to_string(v) +
			                           ": source capacity plus incoming capacities does not fit int32");
		}
	}
	cgp_stats stats;
	std::memset(&stats, 0, sizeof(stats));
	const auto t_start = std::chrono::steady_clock::now();
	mSolved = false;
	const int max_rounds = policy.max_rounds > 0 ? policy.max_rounds : 1 << 20;
	// (a threshold of 2^31 or more closes every arc, as INT32_MAX does)
	int32_t threshold = policy.minimal_residual > 0.0f
	                        ? (int32_t)std::min<double>(std::floor((double)policy.minimal_residual), 2147483647.0)
	                        : 0;

	CGP_HIP(hipMemsetAsync(mFlow.get(), 0, mFlow.bytes(), mStream));
	hipLaunchKernelGGL(k_csr_init_vertices, dim3(grid_blocks(mN)), dim3(256), 0, mStream, mN, mCs.get(), mCt.get(),
	                   mE.get(), mRt.get());
	if (mArcs > 0) {
		hipLaunchKernelGGL(k_csr_copy, dim3(grid_blocks(mArcs)), dim3(256), 0, mStream, mArcs, mArcCap.get(), mRes.get());
	}
	CGP_LAUNCH_CHECK();
	CGP_HIP(hipStreamSynchronize(mStream));
	stats.init_ms = ms_since(t_start);
	trace.emit(CGP_TRACE_STARTED, 0, 0, 0);

	std::vector<int32_t> level_starts;
	int iteration = 0;
	for (;;) {
		if (iteration >= max_rounds) {
			fail(CGP_ERR_NOT_CONVERGED, "max_rounds (" + std::to_string(max_rounds) + ") reached before convergence");
		}
		trace.emit(CGP_TRACE_BEGIN_ITERATION, iteration, 0, 0);
		auto t0 = std::chrono::steady_clock::now();
		bfs(threshold, level_starts, stats);
		stats.relabel_ms += ms_since(t0);
		++stats.rounds;
		trace.emit(CGP_TRACE_AFTER_RELABEL, iteration, 0, 0, level_starts.data(), (int)level_starts.size());
		t0 = std::chrono::steady_clock::now();
		const bool pushed = push_sweep(level_starts, stats);
		stats.discharge_ms += ms_since(t0);
		const bool done = !pushed && threshold == 0;
		if (trace.fn) {
			trace.emit(CGP_TRACE_AFTER_PUSH, iteration, done ? 1 : 0, read_flow());
		}
		++iteration;
		if (!pushed) {
			if (threshold > 0) {
				threshold = 0;  // exact finish
				continue;
			}
			break;
		}
	}
	auto t0 = std::chrono::steady_clock::now();
	source_closure(stats);
	const int64_t flow = read_flow();
	stats.cut_ms = ms_since(t0);
	stats.flow = flow;
	stats.total_ms = ms_since(t_start);
	stats.voxels = mN;
	mSolved = true;
	trace.emit(CGP_TRACE_FINISHED, iteration, 1, flow);
	if (stats_out) {
		*stats_out = stats;
	}
	return flow;
}

void CsrSolver::sink_side(uint8_t *host_out, uint8_t mask, uint8_t background)
{
	if (!mSolved) {
		fail(CGP_ERR_STATE, "fill_segments before max_flow");
	}
	DeviceBuffer<uint8_t> tmp(mN);
	hipLaunchKernelGGL(k_csr_masks, dim3(grid_blocks(mN)), dim3(256), 0, mStream, mN, mH.get(), mS.get(), tmp.get(), 0,
	                   mask, background);
	CGP_LAUNCH_CHECK();
	CGP_HIP(hipMemcpyAsync(host_out, tmp.get(), mN, hipMemcpyDeviceToHost, mStream));
	CGP_HIP(hipStreamSynchronize(mStream));
}

void CsrSolver::source_side(uint8_t *host_out)
{
	if (!mSolved) {
		fail(CGP_ERR_STATE, "source_set before max_flow");
	}
	DeviceBuffer<uint8_t> tmp(mN);
	hipLaunchKernelGGL(k_csr_masks, dim3(grid_blocks(mN)), dim3(256), 0, mStream, mN, mH.get(), mS.get(), tmp.get(), 1,
	                   (uint8_t)1, (uint8_t)0);
	CGP_LAUNCH_CHECK();
	CGP_HIP(hipMemcpyAsync(host_out, tmp.get(), mN, hipMemcpyDeviceToHost, mStream));
	CGP_HIP(hipStreamSynchronize(mStream));
}

void CsrSolver::labels(int32_t *host_out)
{
	if (!mSolved) {
		fail(CGP_ERR_STATE, "labels before max_flow");
	}
	CGP_HIP(hipMemcpyAsync(host_out, mH.get(), mN * sizeof(int32_t), hipMemcpyDeviceToHost, mStream));
	CGP_HIP(hipStreamSynchronize(mStream));
}

}  // namespace cgp
