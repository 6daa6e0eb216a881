// Trace objects, policies and integer flows through the drop-in header, as used by
// benchmarks/graph_cut/graph_cut.cu:82-177, 225-271 (TraceObject, GraphCutPolicy with
// relabelPolicy.edgeTraversalCheck.minimalResidual, max_flow_with_tracing).
// Prints the hook counts and the flow; exit 0 when the flow equals the expected value given on
// the command line and the hooks fired in order.
#include <cugip/graph_cut.hpp>

#include <cstdio>
#include <cstdlib>
#include <vector>

struct CountingTrace {
	int started = 0, iterations = 0, relabels = 0, pushes = 0, finished = 0;
	bool last_done = false;
	long long final_flow = -1;
	template <typename TFlow>
	void computationStarted(const cugip::GraphCutData<TFlow> &) { ++started; }
	void beginIteration(int) { ++iterations; }
	void afterRelabel(int, const std::vector<int> &) { ++relabels; }
	template <typename TFlow>
	void afterPush(int, bool done, const cugip::GraphCutData<TFlow> &) { ++pushes; last_done = done; }
	template <typename TFlow>
	void computationFinished(TFlow flow, const cugip::GraphCutData<TFlow> &) { ++finished; final_flow = (long long)flow; }
};

int main(int argc, char **argv)
{
	// chain s -> 0 -> 1 -> ... -> n-1 -> t with a bottleneck in the middle plus side branches
	const int n = 64;
	cugip::Graph<int32_t> graph;
	graph.set_vertex_count(n);
	std::vector<cugip::EdgeRecord> edges;
	std::vector<int32_t> fwd, bwd;
	for (int i = 0; i + 1 < n; ++i) {
		edges.emplace_back(i, i + 1);
		fwd.push_back(i == n / 2 ? 7 : 40 + i % 5);
		bwd.push_back(3);
	}
	std::vector<int32_t> src(n, 0), snk(n, 0);
	src[0] = 1000;
	src[5] = 11;
	snk[n - 1] = 1000;
	snk[10] = 4;
	graph.set_nweights((int)edges.size(), edges.data(), fwd.data(), bwd.data());
	graph.set_tweights(src.data(), snk.data());
	const long long want = argc > 1 ? std::atoll(argv[1]) : -1;
	int bad = 0;
	for (float threshold : {0.0f, 2.5f}) {
		cugip::GraphCutPolicy policy;
		policy.relabelPolicy.edgeTraversalCheck.minimalResidual = threshold;
		CountingTrace trace;
		const int32_t flow = graph.max_flow_with_tracing(policy, trace);
		std::printf("threshold %.1f flow %d started %d iterations %d relabels %d pushes %d finished %d done %d\n",
		            threshold, flow, trace.started, trace.iterations, trace.relabels, trace.pushes, trace.finished,
		            trace.last_done ? 1 : 0);
		bad |= (want >= 0 && flow != want) || trace.started != 1 || trace.finished != 1 || trace.iterations < 1 ||
		       trace.relabels != trace.iterations || !trace.last_done || trace.final_flow != flow;
	}
	std::vector<uint8_t> smin(n);
	graph.source_set(smin.data());
	int s = 0;
	for (int i = 0; i < n; ++i) {
		s += smin[i];
	}
	std::printf("source_side %d\n", s);
	// invalid input must raise ExceptionBase, not crash
	try {
		cugip::Graph<float> bad_graph;
		bad_graph.set_vertex_count(4);
		const float half[4] = {0.5f, 0, 0, 0}, zero[4] = {0, 0, 0, 0};
		bad_graph.set_tweights(half, zero);
		std::printf("no exception for fractional capacity\n");
		bad = 1;
	} catch (const cugip::ExceptionBase &e) {
		std::printf("fractional capacity rejected: code %d\n", e.code());
	}
	return bad ? 1 : 0;
}
