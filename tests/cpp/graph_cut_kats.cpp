// Known-answer graphs of the reference's own tests, built and solved through the drop-in
// header exactly as the reference's callers do (cugip::Graph<float>, EdgeRecord, set_*,
// max_flow, fill_segments).
//   4x4 grid, test/graph_cut.cu:13-67                      -> max flow 4
//   layered graph, unit_tests/graph_cut_test.cu:22-67,172-197 -> 8192, flipped 16384
//   1000x1000 grid, benchmarks/bench-CUGIP.cu:50-97          -> 1000
// Exit status 0 only if every value matches; prints one line per case.
#include <cugip/graph_cut.hpp>

#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>

static int failures = 0;

static void expect(const char *name, long long got, long long want)
{
	std::printf("%s %lld %s\n", name, got, got == want ? "ok" : "MISMATCH");
	if (got != want) {
		++failures;
	}
}

static void grid_4x4()
{
	cugip::Graph<float> graph;
	graph.set_vertex_count(16);
	const int v1[24] = {0, 1, 2, 4, 5, 6, 8, 9, 10, 12, 13, 14, 0, 1, 2, 3, 4, 5, 6, 7, 8, 9, 10, 11};
	const int v2[24] = {1, 2, 3, 5, 6, 7, 9, 10, 11, 13, 14, 15, 4, 5, 6, 7, 8, 9, 10, 11, 12, 13, 14, 15};
	cugip::EdgeRecord edges[24];
	float weights[24];
	for (int i = 0; i < 24; ++i) {
		edges[i] = cugip::EdgeRecord(v1[i], v2[i]);
		weights[i] = (i >= 16 && i < 20) ? 1.0f : 5.0f;
	}
	float source[16] = {0}, sink[16] = {0};
	for (int i = 0; i < 4; ++i) {
		source[i] = 100.0f;
		sink[12 + i] = 100.0f;
	}
	graph.set_nweights(24, edges, weights, weights);
	graph.set_tweights(source, sink);
	const float flow = graph.max_flow();
	expect("grid4x4", (long long)flow, 4);
	std::vector<uint8_t> seg(16);
	graph.fill_segments(seg.data(), 255, 0);
	int sink_side = 0;
	for (int i = 0; i < 16; ++i) {
		sink_side += seg[i] == 255;
	}
	// rows 2 and 3 reach t after the cut through the weight-1 links
	expect("grid4x4_sink_side", sink_side, 8);
}

static long long layered(bool flip)
{
	const int layers = 100, size = 1024, neighbours = 8;
	cugip::Graph<float> graph;
	graph.set_vertex_count(layers * size);
	std::vector<cugip::EdgeRecord> edges;
	std::vector<float> fwd, bwd;
	for (int layer = 1; layer < layers; ++layer) {
		const float w = layer == layers / 2 ? 1.0f : 20.0f;
		for (int i = 0; i < size; ++i) {
			for (int j = 0; j < neighbours; ++j) {
				edges.emplace_back((layer - 1) * size + i, layer * size + (i + j) % size);
				fwd.push_back(w);
				bwd.push_back(2 * w);
			}
		}
	}
	std::vector<float> ts(layers * size, 0.0f), tt(layers * size, 0.0f);
	for (int i = 0; i < size; ++i) {
		ts[i] = 100000.0f;
		tt[(layers - 1) * size + i] = 100000.0f;
	}
	graph.set_nweights((int)edges.size(), edges.data(), fwd.data(), bwd.data());
	if (flip) {
		graph.set_tweights(tt.data(), ts.data());
	} else {
		graph.set_tweights(ts.data(), tt.data());
	}
	return (long long)graph.max_flow();
}

static void grid_1000()
{
	const int w = 1000, h = 1000;
	std::vector<float> source(w * h), sink(w * h);
	std::vector<cugip::EdgeRecord> edges((w - 1) * h + (h - 1) * w);
	std::vector<float> weights(edges.size()), back(edges.size());
	for (int y = 0; y < h; ++y) {
		for (int x = 0; x < w; ++x) {
			source[x + y * w] = y == 0 ? 100000.0f : 0.0f;
			sink[x + y * w] = y == h - 1 ? 100000.0f : 0.0f;
			if (x < w - 1) {
				edges[x + y * (w - 1)] = cugip::EdgeRecord(x + y * w, (x + 1) + y * w);
				weights[x + y * (w - 1)] = 50;
				back[x + y * (w - 1)] = 50;
			}
			if (y < h - 1) {
				const int e = (w - 1) * h + x + y * w;
				edges[e] = cugip::EdgeRecord(x + y * w, x + (y + 1) * w);
				weights[e] = (float)(std::abs(y - 50) + 1);
				back[e] = (float)(std::abs(y - 50) + 1);
			}
		}
	}
	cugip::Graph<float> graph;
	graph.set_vertex_count(w * h);
	graph.set_nweights((int)edges.size(), edges.data(), weights.data(), back.data());
	graph.set_tweights(source.data(), sink.data());
	expect("grid1000", (long long)graph.max_flow(), 1000);
}

int main(int argc, char **argv)
{
	try {
		grid_4x4();
		expect("layered", layered(false), 8192);
		expect("layered_flipped", layered(true), 16384);
		if (argc < 2 || std::strcmp(argv[1], "--quick") != 0) {
			grid_1000();
		}
	} catch (const cugip::ExceptionBase &e) {
		std::printf("exception code %d: %s\n", e.code(), e.what());
		return 2;
	}
	return failures ? 1 : 0;
}
