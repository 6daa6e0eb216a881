// The batched grid entry of the drop-in header: solves uneven synthetic grids together through
// cugip::GridBatch and one by one through cugip::Graph, and compares flows and S_min bytes.
// usage: graph_cut_batch count conn seed
// Prints one line per grid ("grid i WxHxD flow F ok|MISMATCH"); exit status 0 only if all match.
#include <cugip/graph_cut.hpp>

#include <cstdio>
#include <cstdlib>
#include <vector>

struct Instance {
	int nx, ny, nz;
	std::vector<int32_t> cs, ct, cn;
};

int main(int argc, char **argv)
{
	if (argc < 4) {
		std::fprintf(stderr, "usage: %s count conn seed\n", argv[0]);
		return 2;
	}
	const int count = std::atoi(argv[1]), conn = std::atoi(argv[2]);
	const unsigned long long seed = std::strtoull(argv[3], nullptr, 10);
	std::vector<Instance> grids((size_t)count);
	for (int i = 0; i < count; ++i) {
		Instance &g = grids[(size_t)i];
		// edge lengths 8..71, different per axis and per grid
		g.nx = 8 + (int)((seed * 7 + (unsigned long long)i * 29) % 64);
		g.ny = 8 + (int)((seed * 11 + (unsigned long long)i * 17) % 48);
		g.nz = 8 + (int)((seed * 13 + (unsigned long long)i * 23) % 40);
		const size_t n = (size_t)g.nx * g.ny * g.nz;
		std::vector<uint8_t> volume(n);
		g.cs.resize(n);
		g.ct.resize(n);
		g.cn.resize(n * (size_t)(conn / 2));
		if (cgp_generate_volume(g.nx, g.ny, g.nz, seed, (uint64_t)i, volume.data()) != CGP_OK ||
		    cgp_caps_from_volume(g.nx, g.ny, g.nz, conn, volume.data(), g.cs.data(), g.ct.data(), g.cn.data()) !=
		        CGP_OK) {
			std::fprintf(stderr, "generator failed\n");
			return 2;
		}
	}
	int failures = 0;
	try {
		cugip::GridBatch<int32_t> batch(conn);
		for (const Instance &g : grids) {
			batch.add_grid(g.nx, g.ny, g.nz, g.cs.data(), g.ct.data(), g.cn.data());
		}
		batch.max_flow();
		for (int i = 0; i < count; ++i) {
			const Instance &g = grids[(size_t)i];
			const size_t n = (size_t)g.nx * g.ny * g.nz;
			cugip::Graph<int32_t> single;
			single.set_grid(g.nx, g.ny, g.nz, conn, g.cs.data(), g.ct.data(), g.cn.data());
			single.max_flow();
			std::vector<uint8_t> a(n), b(n);
			batch.source_set(i, a.data());
			single.source_set(b.data());
			const bool same = batch.max_flow_value(i) == single.max_flow_value() && a == b;
			std::printf("grid %d %dx%dx%d flow %lld %s\n", i, g.nx, g.ny, g.nz, (long long)batch.max_flow_value(i),
			            same ? "ok" : "MISMATCH");
			failures += same ? 0 : 1;
		}
	} catch (const cugip::ExceptionBase &e) {
		std::printf("exception code %d: %s\n", e.code(), e.what());
		return 1;
	}
	return failures ? 1 : 0;
}
