// The grid entry of the drop-in header on a synthetic instance: prints the max flow and the
// sizes of S_min and T_min, so the caller can compare them with the oracle.
// usage: graph_cut_grid nx ny nz conn seed
#include <cugip/graph_cut.hpp>

#include <cstdio>
#include <cstdlib>
#include <vector>

int main(int argc, char **argv)
{
	if (argc < 6) {
		std::fprintf(stderr, "usage: %s nx ny nz conn seed\n", argv[0]);
		return 2;
	}
	const int nx = std::atoi(argv[1]), ny = std::atoi(argv[2]), nz = std::atoi(argv[3]), conn = std::atoi(argv[4]);
	const unsigned long long seed = std::strtoull(argv[5], nullptr, 10);
	const size_t n = (size_t)nx * ny * nz;
	std::vector<uint8_t> volume(n);
	std::vector<int32_t> cs(n), ct(n), cn(n * (conn / 2));
	if (cgp_generate_volume(nx, ny, nz, seed, 0, volume.data()) != CGP_OK ||
	    cgp_caps_from_volume(nx, ny, nz, conn, volume.data(), cs.data(), ct.data(), cn.data()) != CGP_OK) {
		std::fprintf(stderr, "generator failed\n");
		return 2;
	}
	try {
		cugip::Graph<int32_t> graph;
		graph.set_grid(nx, ny, nz, conn, cs.data(), ct.data(), cn.data());
		graph.max_flow();
		std::vector<uint8_t> smin(n), tmin(n);
		graph.source_set(smin.data());
		graph.fill_segments(tmin.data(), 1, 0);
		size_t s = 0, t = 0;
		for (size_t i = 0; i < n; ++i) {
			s += smin[i];
			t += tmin[i];
		}
		std::printf("flow %lld source_side %zu sink_side %zu\n", (long long)graph.max_flow_value(), s, t);
	} catch (const cugip::ExceptionBase &e) {
		std::printf("exception code %d: %s\n", e.code(), e.what());
		return 1;
	}
	return 0;
}
