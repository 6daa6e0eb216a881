// Region adjacency graphs in the text format of the reference's region benchmarks
// (benchmarks/graph_cut/graph.hpp:48-104 writes and reads it, benchmarks/graph_cut/
// segmentation_regions.cpp:53-80 reads the marker files), turned into the integer edge list and
// t-links that cgp_set_nweights / cgp_set_tweights take, the way computeCudaGraphCut
// (benchmarks/graph_cut/graph_cut.cpp:100-154) turns them into GraphData<float>.
//
// Graph file, one record per line ('#' lines and empty lines skipped):
//   n;<id>;<count>;<sum>            region <id> (ids 1..n), its voxel count and intensity sum
//   e;<a>;<b>;<count>;<sum>         adjacency of regions a and b with its boundary statistic
// Marker file: "f <id>" (object: sink t-link) or "b <id>" (background: source t-link).
// The edge weight in both directions is the boundary statistic <sum> scaled to an integer:
// round(sum * weight_scale); a marked region gets marker_capacity on its terminal link.
#include <cerrno>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <string>
#include <unordered_map>
#include <vector>

#include "common.h"
#include "solver.h"

namespace cgp {

namespace {

// Fields of a ';'-separated record after its tag ("n;" or "e;"), each trimmed of blanks.
std::vector<std::string> fields(const std::string &line)
{
	std::vector<std::string> out;
	size_t start = 2;
	while (start <= line.size()) {
		size_t end = line.find(';', start);
		if (end == std::string::npos) {
			end = line.size();
		}
		std::string f = line.substr(start, end - start);
		const size_t a = f.find_first_not_of(" \t\r");
		const size_t b = f.find_last_not_of(" \t\r");
		out.push_back(a == std::string::npos ? std::string() : f.substr(a, b - a + 1));
		start = end + 1;
	}
	return out;
}

long long to_int(const std::string &s, const std::string &where)
{
	char *end = nullptr;
	errno = 0;
	const long long v = std::strtoll(s.c_str(), &end, 10);
	if (s.empty() || *end != '\0' || errno) {
		fail(CGP_ERR_INVALID, where + ": '" + s + "' is not an integer");
	}
	return v;
}

double to_real(const std::string &s, const std::string &where)
{
	char *end = nullptr;
	errno = 0;
	const double v = std::strtod(s.c_str(), &end);
	if (s.empty() || *end != '\0' || errno || !std::isfinite(v)) {
		fail(CGP_ERR_INVALID, where + ": '" + s + "' is not a finite number");
	}
	return v;
}

}  // namespace

RegionGraph read_region_graph(const char *path)
{
	std::ifstream file(path);
	if (!file) {
		fail(CGP_ERR_INVALID, std::string(path) + ": cannot open");
	}
	RegionGraph g;
	std::vector<char> seen;
	// the reference keeps its edges in a map keyed by the ordered pair (a, b): a later record of the
	// same pair replaces the earlier one
	std::unordered_map<unsigned long long, size_t> index;
	std::string line;
	long long lineno = 0;
	while (std::getline(file, line)) {
		++lineno;
		if (!line.empty() && line.back() == '\r') {
			line.pop_back();
		}
		if (line.empty() || line[0] == '#') {
			continue;
		}
		const std::string where = std::string(path) + ":" + std::to_string(lineno);
		if (line.size() < 2 || line[1] != ';' || (line[0] != 'n' && line[0] != 'e')) {
			fail(CGP_ERR_INVALID, where + ": unknown record '" + line + "'");
		}
		const std::vector<std::string> f = fields(line);
		if (line[0] == 'n') {
			if (f.size() != 3) {
				fail(CGP_ERR_INVALID, where + ": a node record has 3 fields");
			}
			const long long id = to_int(f[0], where);
			if (id < 1 || id > 0x7fffffffLL) {
				fail(CGP_ERR_INVALID, where + ": node ids start at 1");
			}
			(void)to_int(f[1], where);
			(void)to_real(f[2], where);
			if ((size_t)id > seen.size()) {
				seen.resize((size_t)id, 0);
			}
			seen[(size_t)id - 1] = 1;
		} else {
			if (f.size() != 4) {
				fail(CGP_ERR_INVALID, where + ": an edge record has 4 fields");
			}
			const long long a = to_int(f[0], where), b = to_int(f[1], where);
			if (a < 1 || b < 1 || a > 0x7fffffffLL || b > 0x7fffffffLL) {
				fail(CGP_ERR_INVALID, where + ": region ids start at 1");
			}
			(void)to_int(f[2], where);
			const double sum = to_real(f[3], where);
			if (sum < 0.0) {
				fail(CGP_ERR_INVALID, where + ": negative edge statistic");
			}
			const unsigned long long key = ((unsigned long long)(a - 1) << 32) | (unsigned long long)(b - 1);
			const auto it = index.find(key);
			if (it != index.end()) {
				g.sum[it->second] = sum;
				continue;
			}
			index.emplace(key, g.a.size());
			g.a.push_back((int32_t)(a - 1));
			g.b.push_back((int32_t)(b - 1));
			g.sum.push_back(sum);
		}
	}
	// computeCudaGraphCut takes the node count as the vertex count and indexes by id - 1, so the
	// ids must be exactly 1..n
	g.vertices = (int32_t)seen.size();
	for (size_t i = 0; i < seen.size(); ++i) {
		if (!seen[i]) {
			fail(CGP_ERR_INVALID, std::string(path) + ": node ids must run from 1 to the node count (" +
			                          std::to_string(i + 1) + " is missing)");
		}
	}
	for (size_t i = 0; i < g.a.size(); ++i) {
		if (g.a[i] >= g.vertices || g.b[i] >= g.vertices) {
			fail(CGP_ERR_INVALID, std::string(path) + ": edge " + std::to_string(i) + " names a region without a "
			                          "node record");
		}
	}
	return g;
}

std::vector<std::pair<bool, int32_t>> read_region_markers(const char *path)
{
	std::ifstream file(path);
	if (!file) {
		fail(CGP_ERR_INVALID, std::string(path) + ": cannot open");
	}
	std::vector<std::pair<bool, int32_t>> out;
	std::string line;
	long long lineno = 0;
	while (std::getline(file, line)) {
		++lineno;
		if (!line.empty() && line.back() == '\r') {
			line.pop_back();
		}
		if (line.empty() || line[0] == '#') {
			continue;
		}
		const std::string where = std::string(path) + ":" + std::to_string(lineno);
		if (line[0] != 'f' && line[0] != 'b') {
			fail(CGP_ERR_INVALID, where + ": unknown marker record '" + line + "'");
		}
		std::string rest = line.substr(1);
		const size_t s = rest.find_first_not_of(" \t");
		const size_t e = rest.find_last_not_of(" \t");
		rest = s == std::string::npos ? std::string() : rest.substr(s, e - s + 1);
		const long long id = to_int(rest, where);
		if (id < 1 || id > 0x7fffffffLL) {
			fail(CGP_ERR_INVALID, where + ": region ids start at 1");
		}
		out.emplace_back(line[0] == 'f', (int32_t)(id - 1));
	}
	return out;
}

void region_graph_arrays(const RegionGraph &g, const std::vector<std::pair<bool, int32_t>> &markers,
                         double weight_scale, int32_t marker_capacity, cgp_edge *edges, int32_t *weights,
                         int32_t *cap_source, int32_t *cap_sink)
{
	if (!(weight_scale > 0.0) || !std::isfinite(weight_scale)) {
		fail(CGP_ERR_INVALID, "weight_scale must be positive and finite");
	}
	if (marker_capacity < 0) {
		fail(CGP_ERR_INVALID, "marker_capacity must be non-negative");
	}
	for (size_t i = 0; i < g.a.size(); ++i) {
		const double w = std::nearbyint(g.sum[i] * weight_scale);
		if (w > 1073741823.0) {
			fail(CGP_ERR_OVERFLOW, "edge " + std::to_string(i) + ": scaled weight does not fit (2 x weight < 2^31)");
		}
		edges[i].first = (uint32_t)std::min(g.a[i], g.b[i]);
		edges[i].second = (uint32_t)std::max(g.a[i], g.b[i]);
		weights[i] = (int32_t)w;
	}
	std::memset(cap_source, 0, sizeof(int32_t) * (size_t)g.vertices);
	std::memset(cap_sink, 0, sizeof(int32_t) * (size_t)g.vertices);
	for (const auto &m : markers) {
		if (m.second >= g.vertices) {
			fail(CGP_ERR_INVALID, "marker of region " + std::to_string(m.second + 1) + " outside the graph");
		}
		// graph_cut.cpp:128: background markers get the source link, object markers the sink link
		cap_source[m.second] = m.first ? 0 : marker_capacity;
		cap_sink[m.second] = m.first ? marker_capacity : 0;
	}
}

}  // namespace cgp
