// cugip/graph_cut.hpp -- header-only C++ drop-in for the reference's graph cut API.
//
// Mirrors include/cugip/advanced_operations/graph_cut.hpp (Graph<TFlow>, DummyTraceObject),
// detail/edge_record.hpp (EdgeRecord, EdgeReverseTraversable, EdgeForwardTraversable),
// detail/graph_cut_policies.hpp (RelabelPolicy, PushPolicy, GraphCutPolicy) and
// detail/graph_cut_host_utils.hpp (GraphData) of JanKolomaznik/cugip, so that callers such as
// test/graph_cut.cu and benchmarks/graph_cut/graph_cut.cu compile against it after replacing the
// include path. Every call forwards to the C ABI of libcugip_hip.so (cugip_hip.h), which runs the
// max flow on an AMD Instinct MI355X with hand-written HIP kernels.
//
// Differences a caller can observe, all deliberate:
//  * flows are exact integers. TFlow may be int32_t (or int64_t) or float; float capacities must
//    be whole numbers that fit int32, otherwise ExceptionBase is thrown. max_flow() returns TFlow;
//    max_flow_value() returns the exact int64 value (the total can exceed 2^31 and 2^24);
//  * sink t-links are finite capacities (the reference drains any excess into t once the sink
//    capacity is positive, graph_cut_push.hpp:26-41);
//  * fill_segments marks the vertices that still reach t, as the reference does; source_set()
//    additionally returns the canonical minimum cut S_min;
//  * trace hooks receive std::vector<int> level starts instead of thrust::host_vector<int>, and a
//    host-side GraphCutData snapshot instead of device pointers;
//  * set_grid() is new: 3D grid graphs with implicit 6/26-connected edges (no CSR);
//  * GridBatch is new: many independent grids solved together.
#pragma once

#include <cstdint>
#include <cmath>
#include <exception>
#include <iostream>
#include <limits>
#include <memory>
#include <string>
#include <type_traits>
#include <vector>

#include "../cugip_hip.h"

namespace cugip {

// ---- errors (exception.hpp:20-46) -------------------------------------------------------------
class ExceptionBase : public std::exception {
public:
	ExceptionBase(int aCode, std::string aMessage) : mCode(aCode), mMessage(std::move(aMessage)) {}
	const char *what() const noexcept override { return mMessage.c_str(); }
	int code() const noexcept { return mCode; }

private:
	int mCode;
	std::string mMessage;
};

namespace detail {
inline void check(cgp_graph aGraph, int aStatus)
{
	if (aStatus != CGP_OK) {
		throw ExceptionBase(aStatus, cgp_last_error(aGraph));
	}
}

template <typename TFlow>
inline int32_t to_capacity(TFlow aValue)
{
	if (std::is_floating_point<TFlow>::value) {
		const double v = static_cast<double>(aValue);
		if (!(v >= 0.0) || v > 2147483647.0 || std::floor(v) != v) {
			throw ExceptionBase(CGP_ERR_INVALID, "capacities must be non-negative whole numbers below 2^31");
		}
		return static_cast<int32_t>(v);
	}
	const long long v = static_cast<long long>(aValue);
	if (v < 0 || v > 2147483647LL) {
		throw ExceptionBase(CGP_ERR_INVALID, "capacities must be non-negative and below 2^31");
	}
	return static_cast<int32_t>(v);
}

template <typename TFlow>
inline std::vector<int32_t> to_capacities(const TFlow *aValues, size_t aCount)
{
	std::vector<int32_t> out(aCount);
	for (size_t i = 0; i < aCount; ++i) {
		out[i] = to_capacity(aValues[i]);
	}
	return out;
}

template <typename TFlow>
inline TFlow narrow_flow(int64_t aFlow)
{
	if (!std::is_floating_point<TFlow>::value &&
	    aFlow > static_cast<int64_t>(std::numeric_limits<TFlow>::max())) {
		throw ExceptionBase(CGP_ERR_OVERFLOW, "max-flow value does not fit TFlow; use max_flow_value()");
	}
	return static_cast<TFlow>(aFlow);
}
}  // namespace detail

// ---- edges (edge_record.hpp:8-83) -------------------------------------------------------------
typedef unsigned NodeId;
typedef unsigned long CombinedNodeId;

struct EdgeRecord {
	EdgeRecord(NodeId aFirst, NodeId aSecond)
	{
		first = aFirst < aSecond ? aFirst : aSecond;
		second = aFirst < aSecond ? aSecond : aFirst;
	}
	EdgeRecord() : edgeCombIdx(0) {}

	union {
		CombinedNodeId edgeCombIdx;
		struct {
			NodeId second;
			NodeId first;
		};
	};
};
static_assert(sizeof(EdgeRecord) == sizeof(cgp_edge), "EdgeRecord must match cgp_edge");

struct EdgeReverseTraversable {
	EdgeReverseTraversable() : minimalResidual(0.0f) {}
	float minimalResidual;
};

struct EdgeForwardTraversable {};

// ---- policies (graph_cut_policies.hpp:7-89) ---------------------------------------------------
enum class RelabelImplementation { Default, Naive, OptimizedNaive };
enum class PreflowInitialization { Default, Push };
enum class TLinkType : int { Source = 0, Sink = 1 };

template <int tThreadCount = 512, int tGranularity = 64, typename TEdgeCheck = EdgeReverseTraversable,
          TLinkType tStartTLinkType = TLinkType::Sink>
struct RelabelPolicy {
	static constexpr RelabelImplementation cRelabelImplementation = RelabelImplementation::Naive;
	static constexpr TLinkType cStartTLinkType = tStartTLinkType;
	enum {
		INVALID_LABEL = CGP_LABEL_UNREACHABLE,
		THREADS = tThreadCount,
		SCHEDULE_GRANULARITY = tGranularity,
	};
	RelabelPolicy(int aMaxLevels = (1 << 30)) : maxAssignedLevels(aMaxLevels) {}
	int maxLevels() const { return maxAssignedLevels; }
	TEdgeCheck edgeTraversalCheck;
	int maxAssignedLevels;
};

template <int tThreadCount = 512>
struct PushPolicy {
	enum { THREADS = tThreadCount };
};

struct GraphCutPolicy {
	static constexpr PreflowInitialization cPreflowInitialization = PreflowInitialization::Default;
	typedef cugip::RelabelPolicy<512, 64, EdgeReverseTraversable> RelabelPolicy;
	typedef cugip::PushPolicy<512> PushPolicy;

	RelabelPolicy relabelPolicy;

	// MI355X solver schedule (0 = library default); see cgp_policy in cugip_hip.h
	int tileX = 0, tileY = 0, tileZ = 0;
	int dischargesPerRelabel = 0;
	int maxTileIterations = 0;
	int maxRounds = 0;
	int engine = CGP_ENGINE_AUTO;
};

// ---- host-side view handed to trace objects (graph_cut_data.hpp:20-140) ------------------------
template <typename TFlow>
struct GraphCutData {
	typedef TFlow Flow;
	int vertexCount() const { return mVertexCount; }
	int mVertexCount = 0;
	int64_t mFlow = 0;       // flow into t so far
	cgp_stats mStats{};      // filled for computationFinished
};

struct DummyTraceObject {
	template <typename TFlow>
	void computationStarted(const GraphCutData<TFlow> &) {}
	void beginIteration(int) {}
	void afterRelabel(int, const std::vector<int> &) {}
	template <typename TFlow>
	void afterPush(int, bool, const GraphCutData<TFlow> &) {}
	template <typename TFlow>
	void computationFinished(TFlow, const GraphCutData<TFlow> &) {}
};

// ---- host graph builder (graph_cut_host_utils.hpp:8-48) ---------------------------------------
template <typename TWeight>
struct GraphData {
	void setVertexCount(int aVertexCount)
	{
		tlinksSource.resize(aVertexCount);
		tlinksSink.resize(aVertexCount);
	}
	void setTWeights(int aVertex, TWeight aSourceWeight, TWeight aSinkWeight)
	{
		tlinksSource[aVertex] = aSourceWeight;
		tlinksSink[aVertex] = aSinkWeight;
	}
	void addEdge(int aVertex1, int aVertex2, TWeight aForwardWeight, TWeight aBackwardWeight)
	{
		edges.emplace_back(aVertex1, aVertex2);
		weights.push_back(aForwardWeight);
		weightsBackward.push_back(aBackwardWeight);
	}
	void reserve(int aVertexCount, int aEdgeCount)
	{
		tlinksSource.reserve(aVertexCount);
		tlinksSink.reserve(aVertexCount);
		edges.reserve(aEdgeCount);
		weights.reserve(aEdgeCount);
		weightsBackward.reserve(aEdgeCount);
	}
	std::vector<TWeight> tlinksSource;
	std::vector<TWeight> tlinksSink;
	std::vector<EdgeRecord> edges;
	std::vector<TWeight> weights;
	std::vector<TWeight> weightsBackward;
};

// ---- the solver (graph_cut.hpp:47-146) --------------------------------------------------------
template <typename TFlow>
class Graph {
public:
	typedef int VertexId;
	typedef TFlow EdgeWeight;

	Graph()
	{
		cgp_graph g = nullptr;
		detail::check(nullptr, cgp_create(&g));
		mHandle.reset(g);
	}

	void set_vertex_count(int aCount)
	{
		detail::check(handle(), cgp_set_vertex_count(handle(), aCount));
		mVertexCount = aCount;
	}

	void set_nweights(int aEdgeCount, const EdgeRecord *aEdges, const EdgeWeight *aWeightsForward,
	                  const EdgeWeight *aWeightsBackward)
	{
		const std::vector<int32_t> fwd = detail::to_capacities(aWeightsForward, (size_t)aEdgeCount);
		const std::vector<int32_t> bwd = detail::to_capacities(aWeightsBackward, (size_t)aEdgeCount);
		detail::check(handle(), cgp_set_nweights(handle(), aEdgeCount, reinterpret_cast<const cgp_edge *>(aEdges),
		                                         fwd.data(), bwd.data()));
	}

	void set_tweights(const EdgeWeight *aCapSource, const EdgeWeight *aCapSink)
	{
		const std::vector<int32_t> src = detail::to_capacities(aCapSource, (size_t)mVertexCount);
		const std::vector<int32_t> snk = detail::to_capacities(aCapSink, (size_t)mVertexCount);
		detail::check(handle(), cgp_set_tweights(handle(), src.data(), snk.data()));
	}

	// 3D grid with implicit edges; see cgp_set_grid for the layout of aCapN / aCapNBack.
	void set_grid(int aNx, int aNy, int aNz, int aConnectivity, const EdgeWeight *aCapSource,
	              const EdgeWeight *aCapSink, const EdgeWeight *aCapN, const EdgeWeight *aCapNBack = nullptr)
	{
		const size_t n = (size_t)aNx * aNy * aNz;
		const size_t planes = (size_t)(aConnectivity / 2) * n;
		const std::vector<int32_t> src = detail::to_capacities(aCapSource, n);
		const std::vector<int32_t> snk = detail::to_capacities(aCapSink, n);
		const std::vector<int32_t> cn = detail::to_capacities(aCapN, planes);
		std::vector<int32_t> cb;
		if (aCapNBack) {
			cb = detail::to_capacities(aCapNBack, planes);
		}
		detail::check(handle(), cgp_set_grid(handle(), aNx, aNy, aNz, aConnectivity, src.data(), snk.data(), cn.data(),
		                                     aCapNBack ? cb.data() : nullptr));
		mVertexCount = (int)n;
	}

	// Loads an MFI instance file (the grid benchmark format of benchmarks/mfi.h); returns its header.
	cgp_mfi_info set_grid_mfi(const std::string &aPath)
	{
		cgp_mfi_info info;
		detail::check(handle(), cgp_set_grid_mfi(handle(), aPath.c_str(), &info));
		mVertexCount = info.width * info.height * info.depth;
		return info;
	}

	TFlow max_flow()
	{
		DummyTraceObject dummyTraceObject;
		return max_flow_with_tracing(dummyTraceObject);
	}

	template <typename TPolicy>
	TFlow max_flow(TPolicy &aPolicy)
	{
		DummyTraceObject dummyTraceObject;
		return max_flow_with_tracing(aPolicy, dummyTraceObject);
	}

	template <typename TTraceObject>
	TFlow max_flow_with_tracing(TTraceObject &aTraceObject)
	{
		GraphCutPolicy policy;
		return max_flow_with_tracing(policy, aTraceObject);
	}

	template <typename TPolicy, typename TTraceObject>
	TFlow max_flow_with_tracing(TPolicy &aPolicy, TTraceObject &aTraceObject)
	{
		apply_policy(aPolicy);
		TraceContext<TTraceObject> context{&aTraceObject, GraphCutData<TFlow>()};
		context.data.mVertexCount = mVertexCount;
		detail::check(handle(), cgp_set_trace(handle(), &TraceContext<TTraceObject>::callback, &context));
		int64_t flow = 0;
		const int status = cgp_max_flow(handle(), &flow, &mStats);
		cgp_set_trace(handle(), nullptr, nullptr);
		detail::check(handle(), status);
		mFlow = flow;
		context.data.mFlow = flow;
		context.data.mStats = mStats;
		const TFlow result = detail::narrow_flow<TFlow>(flow);
		aTraceObject.computationFinished(result, context.data);
		return result;
	}

	// Exact max-flow value of the last solve.
	int64_t max_flow_value() const { return mFlow; }
	const cgp_stats &stats() const { return mStats; }

	template <typename THostArrayView>
	void fill_segments(THostArrayView aVertices, uint8_t mMask = 255, uint8_t mBackground = 0) const
	{
		std::vector<uint8_t> tmp((size_t)mVertexCount);
		detail::check(handle(), cgp_fill_segments(handle(), tmp.data(), mMask, mBackground));
		for (int i = 0; i < mVertexCount; ++i) {
			aVertices[i] = tmp[i];
		}
	}

	// Canonical minimum cut S_min: 1 where the vertex is reachable from s in the final residual graph.
	template <typename THostArrayView>
	void source_set(THostArrayView aVertices) const
	{
		std::vector<uint8_t> tmp((size_t)mVertexCount);
		detail::check(handle(), cgp_source_set(handle(), tmp.data()));
		for (int i = 0; i < mVertexCount; ++i) {
			aVertices[i] = tmp[i];
		}
	}

	void debug_print()
	{
		std::vector<int32_t> labels((size_t)mVertexCount);
		detail::check(handle(), cgp_labels(handle(), labels.data()));
		for (int i = 0; i < mVertexCount; ++i) {
			std::cout << i << ": " << labels[i] << "\n";
		}
	}

	cgp_graph handle() const { return mHandle.get(); }

private:
	struct Deleter {
		void operator()(cgp_graph g) const { cgp_destroy(g); }
	};

	template <typename TTraceObject>
	struct TraceContext {
		TTraceObject *object;
		GraphCutData<TFlow> data;

		static void callback(void *aUser, int32_t aEvent, int32_t aIteration, int32_t aDone, int64_t aFlow,
		                     const int32_t *aLevels, int32_t aLevelCount)
		{
			TraceContext *self = static_cast<TraceContext *>(aUser);
			self->data.mFlow = aFlow;
			switch (aEvent) {
			case CGP_TRACE_STARTED:
				self->object->computationStarted(self->data);
				break;
			case CGP_TRACE_BEGIN_ITERATION:
				self->object->beginIteration(aIteration);
				break;
			case CGP_TRACE_AFTER_RELABEL: {
				std::vector<int> starts;
				if (aLevels && aLevelCount > 0) {
					starts.assign(aLevels, aLevels + aLevelCount);
				}
				self->object->afterRelabel(aIteration, starts);
				break;
			}
			case CGP_TRACE_AFTER_PUSH:
				self->object->afterPush(aIteration, aDone != 0, self->data);
				break;
			default:
				break;  // computationFinished is called after the solve with the final value
			}
		}
	};

	template <typename TPolicy>
	void apply_policy(const TPolicy &aPolicy)
	{
		cgp_policy p;
		detail::check(handle(), cgp_get_policy(handle(), &p));
		p.minimal_residual = aPolicy.relabelPolicy.edgeTraversalCheck.minimalResidual;
		apply_schedule(p, aPolicy, 0);
		detail::check(handle(), cgp_set_policy(handle(), &p));
	}

	template <typename TPolicy>
	static auto apply_schedule(cgp_policy &p, const TPolicy &aPolicy, int) -> decltype(aPolicy.tileX, void())
	{
		p.tile_x = aPolicy.tileX;
		p.tile_y = aPolicy.tileY;
		p.tile_z = aPolicy.tileZ;
		p.discharges_per_relabel = aPolicy.dischargesPerRelabel;
		p.max_tile_iterations = aPolicy.maxTileIterations;
		p.max_rounds = aPolicy.maxRounds;
		p.engine = aPolicy.engine;
	}
	template <typename TPolicy>
	static void apply_schedule(cgp_policy &, const TPolicy &, long)
	{
	}

	std::unique_ptr<cgp_graph_s, Deleter> mHandle;
	int mVertexCount = 0;
	int64_t mFlow = 0;
	cgp_stats mStats{};
};

// Independent 3D grids of one connectivity, solved together in the same launches
// (cgp_set_grid_batch / cgp_solve_batch). No counterpart in the reference, whose Graph<TFlow>
// holds one graph; each grid has the layout of Graph<TFlow>::set_grid.
template <typename TFlow>
class GridBatch {
public:
	explicit GridBatch(int aConnectivity = 6) : mConnectivity(aConnectivity)
	{
		cgp_graph g = nullptr;
		detail::check(nullptr, cgp_create(&g));
		mHandle.reset(g);
	}

	// Copies the capacities; returns the grid's index in the batch.
	int add_grid(int aNx, int aNy, int aNz, const TFlow *aCapSource, const TFlow *aCapSink, const TFlow *aCapN,
	             const TFlow *aCapNBack = nullptr)
	{
		const size_t n = (size_t)aNx * aNy * aNz;
		const size_t planes = (size_t)(mConnectivity / 2) * n;
		Grid grid;
		grid.nx = aNx;
		grid.ny = aNy;
		grid.nz = aNz;
		grid.source = detail::to_capacities(aCapSource, n);
		grid.sink = detail::to_capacities(aCapSink, n);
		grid.n = detail::to_capacities(aCapN, planes);
		if (aCapNBack) {
			grid.back = detail::to_capacities(aCapNBack, planes);
		}
		mGrids.push_back(std::move(grid));
		mUploaded = false;
		return (int)mGrids.size() - 1;
	}

	int size() const { return (int)mGrids.size(); }

	// Solves every grid; returns the per-grid max-flow values.
	std::vector<TFlow> max_flow()
	{
		GraphCutPolicy policy;
		return max_flow(policy);
	}

	template <typename TPolicy>
	std::vector<TFlow> max_flow(TPolicy &aPolicy)
	{
		upload();
		cgp_policy p;
		detail::check(handle(), cgp_get_policy(handle(), &p));
		p.minimal_residual = aPolicy.relabelPolicy.edgeTraversalCheck.minimalResidual;
		p.tile_x = aPolicy.tileX;
		p.tile_y = aPolicy.tileY;
		p.tile_z = aPolicy.tileZ;
		p.discharges_per_relabel = aPolicy.dischargesPerRelabel;
		p.max_tile_iterations = aPolicy.maxTileIterations;
		p.max_rounds = aPolicy.maxRounds;
		p.engine = aPolicy.engine;
		detail::check(handle(), cgp_set_policy(handle(), &p));
		mFlows.assign(mGrids.size(), 0);
		detail::check(handle(), cgp_solve_batch(handle(), mFlows.data(), &mStats));
		std::vector<TFlow> result;
		result.reserve(mFlows.size());
		for (int64_t f : mFlows) {
			result.push_back(detail::narrow_flow<TFlow>(f));
		}
		return result;
	}

	int64_t max_flow_value(int aIndex) const { return mFlows.at((size_t)aIndex); }
	const cgp_stats &stats() const { return mStats; }

	// Canonical minimum cut S_min of grid aIndex: 1 where the voxel is reachable from s.
	template <typename THostArrayView>
	void source_set(int aIndex, THostArrayView aVertices) const
	{
		std::vector<uint8_t> tmp(voxels(aIndex));
		detail::check(handle(), cgp_batch_source_set(handle(), aIndex, tmp.data()));
		for (size_t i = 0; i < tmp.size(); ++i) {
			aVertices[i] = tmp[i];
		}
	}

	// fill_segments of grid aIndex (voxels that still reach t).
	template <typename THostArrayView>
	void fill_segments(int aIndex, THostArrayView aVertices, uint8_t mMask = 255, uint8_t mBackground = 0) const
	{
		std::vector<uint8_t> tmp(voxels(aIndex));
		detail::check(handle(), cgp_batch_fill_segments(handle(), aIndex, tmp.data(), mMask, mBackground));
		for (size_t i = 0; i < tmp.size(); ++i) {
			aVertices[i] = tmp[i];
		}
	}

	cgp_graph handle() const { return mHandle.get(); }

private:
	struct Deleter {
		void operator()(cgp_graph g) const { cgp_destroy(g); }
	};
	struct Grid {
		int nx = 0, ny = 0, nz = 0;
		std::vector<int32_t> source, sink, n, back;
	};

	size_t voxels(int aIndex) const
	{
		const Grid &g = mGrids.at((size_t)aIndex);
		return (size_t)g.nx * g.ny * g.nz;
	}

	void upload()
	{
		if (mUploaded) {
			return;
		}
		std::vector<cgp_grid_instance> desc(mGrids.size());
		for (size_t i = 0; i < mGrids.size(); ++i) {
			const Grid &g = mGrids[i];
			desc[i].nx = g.nx;
			desc[i].ny = g.ny;
			desc[i].nz = g.nz;
			desc[i].reserved0 = 0;
			desc[i].cap_source = g.source.data();
			desc[i].cap_sink = g.sink.data();
			desc[i].cap_n = g.n.data();
			desc[i].cap_n_back = g.back.empty() ? nullptr : g.back.data();
		}
		detail::check(handle(), cgp_set_grid_batch(handle(), (int32_t)desc.size(), mConnectivity, desc.data(), 0));
		mUploaded = true;
	}

	std::unique_ptr<cgp_graph_s, Deleter> mHandle;
	int mConnectivity;
	std::vector<Grid> mGrids;
	bool mUploaded = false;
	std::vector<int64_t> mFlows;
	cgp_stats mStats{};
};

}  // namespace cugip
