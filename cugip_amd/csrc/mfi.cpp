// MFI instance files (benchmarks/mfi.cpp:105-200 of the reference reads them): the packed header,
// source and sink capacity planes, one capacity plane per arc direction, and the active mask.
// Conversion to the grid layout of cgp_set_grid, and a writer for 3D grids.
#include "mfi.h"

#include <cstdio>
#include <cstring>
#include <memory>

#include "common.h"
#include "grid_dirs.h"

namespace cgp {

namespace {

#pragma pack(push, 1)
struct Header {
	int32_t dimension, connectivity;
	int32_t width, height, depth;
	int64_t maxflow;
	int32_t symmetric_arcs;
	int8_t metadata[4096];
	int32_t active_count;
	int32_t type_terminal_cap, type_neighbor_cap;
};
#pragma pack(pop)
static_assert(sizeof(Header) == 5 * 4 + 8 + 4 + 4096 + 3 * 4, "MFI header is packed");

// Arc offsets (dx, dy, dz) in the order of the format's arc enumeration (benchmarks/mfi.h:33-83):
// 2D: 4-connected arcs 0-3, 8-connected adds 4-7; 3D: 6-connected 0-5, 18-connected adds 6-17,
// 26-connected adds 18-25.
const int8_t kArcs2D[8][3] = {{-1, 0, 0}, {1, 0, 0}, {0, -1, 0}, {0, 1, 0},
                              {-1, -1, 0}, {1, -1, 0}, {-1, 1, 0}, {1, 1, 0}};
const int8_t kArcs3D[26][3] = {
	{-1, 0, 0},  {1, 0, 0},  {0, -1, 0},  {0, 1, 0},  {0, 0, -1}, {0, 0, 1},  {-1, -1, 0},
	{1, -1, 0},  {-1, 1, 0}, {1, 1, 0},   {0, -1, -1}, {0, 1, -1}, {0, -1, 1}, {0, 1, 1},
	{-1, 0, -1}, {-1, 0, 1}, {1, 0, -1},  {1, 0, 1},   {-1, -1, -1}, {1, -1, -1}, {-1, 1, -1},
	{1, 1, -1},  {-1, -1, 1}, {1, -1, 1}, {-1, 1, 1},  {1, 1, 1}};

struct FileCloser {
	void operator()(FILE *f) const { std::fclose(f); }
};
using File = std::unique_ptr<FILE, FileCloser>;

int type_size(int32_t type)
{
	switch (type) {
	case 0:
		return 1;
	case 1:
		return 2;
	case 2:
		return 4;
	default:
		return 0;
	}
}

void read_plane(FILE *f, int32_t type, size_t n, std::vector<uint32_t> &out, const std::string &path)
{
	const int size = type_size(type);
	std::vector<uint8_t> raw(n * (size_t)size);
	if (n && std::fread(raw.data(), raw.size(), 1, f) != 1) {
		fail(CGP_ERR_INVALID, path + ": truncated MFI file");
	}
	out.resize(n);
	for (size_t i = 0; i < n; ++i) {
		uint32_t v = 0;
		std::memcpy(&v, raw.data() + i * (size_t)size, (size_t)size);  // little endian
		out[i] = v;
	}
}

int32_t to_capacity(uint32_t v)
{
	if (v > 0x7fffffffu) {
		fail(CGP_ERR_OVERFLOW, "an MFI capacity exceeds 2^31 - 1");
	}
	return (int32_t)v;
}

int find_arc(const int8_t (*arcs)[3], int count, int dx, int dy, int dz)
{
	for (int a = 0; a < count; ++a) {
		if (arcs[a][0] == dx && arcs[a][1] == dy && arcs[a][2] == dz) {
			return a;
		}
	}
	return -1;
}

}  // namespace

MfiFile mfi_read(const std::string &path)
{
	File f(std::fopen(path.c_str(), "rb"));
	if (!f) {
		fail(CGP_ERR_INVALID, path + ": cannot open");
	}
	char magic[3];
	Header h;
	if (std::fread(magic, 3, 1, f.get()) != 1 || std::memcmp(magic, "MFI", 3) != 0) {
		fail(CGP_ERR_INVALID, path + ": not an MFI file");
	}
	if (std::fread(&h, sizeof(h), 1, f.get()) != 1) {
		fail(CGP_ERR_INVALID, path + ": truncated MFI header");
	}
	const bool valid = (h.dimension == 2 && (h.connectivity == 4 || h.connectivity == 8)) ||
	                   (h.dimension == 3 && (h.connectivity == 6 || h.connectivity == 18 || h.connectivity == 26));
	if (!valid) {
		fail(CGP_ERR_INVALID, path + ": unsupported MFI dimension/connectivity " + std::to_string(h.dimension) +
		                          "/" + std::to_string(h.connectivity));
	}
	if (!type_size(h.type_terminal_cap) || !type_size(h.type_neighbor_cap)) {
		fail(CGP_ERR_INVALID, path + ": unknown MFI capacity type");
	}
	const int depth = h.dimension == 2 ? 1 : h.depth;
	if (h.width <= 0 || h.height <= 0 || depth <= 0 ||
	    (int64_t)h.width * h.height * depth >= (int64_t)1 << 31) {
		fail(CGP_ERR_INVALID, path + ": invalid MFI dimensions");
	}
	MfiFile out;
	std::memset(&out.info, 0, sizeof(out.info));
	cgp_mfi_info &info = out.info;
	info.dimension = h.dimension;
	info.connectivity = h.connectivity;
	info.width = h.width;
	info.height = h.height;
	info.depth = depth;
	info.symmetric_arcs = h.symmetric_arcs;
	info.maxflow = h.maxflow;
	info.active_count = h.active_count;
	info.type_terminal_cap = h.type_terminal_cap;
	info.type_neighbor_cap = h.type_neighbor_cap;
	info.grid_conn = (h.connectivity == 4 || h.connectivity == 6) ? 6 : 26;
	std::memcpy(info.metadata, h.metadata, sizeof(info.metadata));
	info.metadata[sizeof(info.metadata) - 1] = 0;
	const size_t n = (size_t)h.width * h.height * depth;
	read_plane(f.get(), h.type_terminal_cap, n, out.source, path);
	read_plane(f.get(), h.type_terminal_cap, n, out.sink, path);
	out.arcs.resize((size_t)h.connectivity);
	for (int a = 0; a < h.connectivity; ++a) {
		read_plane(f.get(), h.type_neighbor_cap, n, out.arcs[(size_t)a], path);
	}
	out.active.resize(n);
	if (n && std::fread(out.active.data(), n, 1, f.get()) != 1) {
		fail(CGP_ERR_INVALID, path + ": truncated MFI active mask");
	}
	return out;
}

MfiGrid mfi_to_grid(const MfiFile &file)
{
	const cgp_mfi_info &info = file.info;
	MfiGrid g;
	g.nx = info.width;
	g.ny = info.height;
	g.nz = info.depth;
	g.conn = info.grid_conn;
	const int half = g.conn / 2;
	const size_t n = (size_t)g.nx * g.ny * g.nz;
	const int8_t(*arcs)[3] = info.dimension == 2 ? kArcs2D : kArcs3D;
	const int8_t(*halves)[3] = g.conn == 6 ? CGP_HALF_OFFSETS_6 : CGP_HALF_OFFSETS_26;
	g.cs.resize(n);
	g.ct.resize(n);
	for (size_t v = 0; v < n; ++v) {
		g.cs[v] = to_capacity(file.source[v]);
		g.ct[v] = to_capacity(file.sink[v]);
	}
	g.cn.assign(n * (size_t)half, 0);
	g.cnb.assign(n * (size_t)half, 0);
	for (int j = 0; j < half; ++j) {
		const int dx = halves[j][0], dy = halves[j][1], dz = halves[j][2];
		// the arc v -> v + o_j, and the arc of the opposite direction leaving v + o_j (absent arcs
		// of a lower connectivity, such as the in-plane diagonals of a 4-connected image, stay 0)
		const int fwd = find_arc(arcs, info.connectivity, dx, dy, dz);
		const int bwd = find_arc(arcs, info.connectivity, -dx, -dy, -dz);
		if (fwd < 0 && bwd < 0) {
			continue;
		}
		for (int z = 0; z < g.nz; ++z) {
			for (int y = 0; y < g.ny; ++y) {
				for (int x = 0; x < g.nx; ++x) {
					const int ux = x + dx, uy = y + dy, uz = z + dz;
					if (ux < 0 || ux >= g.nx || uy < 0 || uy >= g.ny || uz < 0 || uz >= g.nz) {
						continue;
					}
					const size_t v = ((size_t)z * g.ny + y) * g.nx + x;
					const size_t u = ((size_t)uz * g.ny + uy) * g.nx + ux;
					if (fwd >= 0) {
						g.cn[(size_t)j * n + v] = to_capacity(file.arcs[(size_t)fwd][v]);
					}
					if (bwd >= 0) {
						g.cnb[(size_t)j * n + v] = to_capacity(file.arcs[(size_t)bwd][u]);
					}
				}
			}
		}
	}
	return g;
}

void mfi_write_grid(const std::string &path, int nx, int ny, int nz, int conn, const int32_t *cs, const int32_t *ct,
                    const int32_t *cn, const int32_t *cnb, int64_t maxflow, const char *metadata)
{
	if (nx <= 0 || ny <= 0 || nz <= 0 || (conn != 6 && conn != 26) || !cs || !ct || !cn) {
		fail(CGP_ERR_INVALID, "invalid grid for an MFI file");
	}
	const size_t n = (size_t)nx * ny * nz;
	const int half = conn / 2;
	const int8_t(*halves)[3] = conn == 6 ? CGP_HALF_OFFSETS_6 : CGP_HALF_OFFSETS_26;
	std::vector<std::vector<uint32_t>> arcs((size_t)conn, std::vector<uint32_t>(n, 0));
	auto check = [](int32_t v) {
		if (v < 0) {
			fail(CGP_ERR_INVALID, "capacities must be non-negative");
		}
		return (uint32_t)v;
	};
	for (int j = 0; j < half; ++j) {
		const int dx = halves[j][0], dy = halves[j][1], dz = halves[j][2];
		const int fwd = find_arc(kArcs3D, conn, dx, dy, dz);
		const int bwd = find_arc(kArcs3D, conn, -dx, -dy, -dz);
		for (int z = 0; z < nz; ++z) {
			for (int y = 0; y < ny; ++y) {
				for (int x = 0; x < nx; ++x) {
					const int ux = x + dx, uy = y + dy, uz = z + dz;
					if (ux < 0 || ux >= nx || uy < 0 || uy >= ny || uz < 0 || uz >= nz) {
						continue;
					}
					const size_t v = ((size_t)z * ny + y) * nx + x;
					const size_t u = ((size_t)uz * ny + uy) * nx + ux;
					arcs[(size_t)fwd][v] = check(cn[(size_t)j * n + v]);
					arcs[(size_t)bwd][u] = check(cnb ? cnb[(size_t)j * n + v] : cn[(size_t)j * n + v]);
				}
			}
		}
	}
	Header h;
	std::memset(&h, 0, sizeof(h));
	h.dimension = 3;
	h.connectivity = conn;
	h.width = nx;
	h.height = ny;
	h.depth = nz;
	h.maxflow = maxflow;
	h.symmetric_arcs = cnb ? 0 : 1;
	if (metadata) {
		std::strncpy(reinterpret_cast<char *>(h.metadata), metadata, sizeof(h.metadata) - 1);
	}
	h.active_count = 0;
	h.type_terminal_cap = 2;
	h.type_neighbor_cap = 2;
	File f(std::fopen(path.c_str(), "wb"));
	if (!f) {
		fail(CGP_ERR_INVALID, path + ": cannot create");
	}
	bool ok = std::fwrite("MFI", 3, 1, f.get()) == 1 && std::fwrite(&h, sizeof(h), 1, f.get()) == 1;
	std::vector<uint32_t> plane(n);
	for (size_t v = 0; ok && v < n; ++v) {
		plane[v] = check(cs[v]);
	}
	ok = ok && std::fwrite(plane.data(), sizeof(uint32_t), n, f.get()) == n;
	for (size_t v = 0; ok && v < n; ++v) {
		plane[v] = check(ct[v]);
	}
	ok = ok && std::fwrite(plane.data(), sizeof(uint32_t), n, f.get()) == n;
	for (int a = 0; ok && a < conn; ++a) {
		ok = std::fwrite(arcs[(size_t)a].data(), sizeof(uint32_t), n, f.get()) == n;
	}
	const std::vector<uint8_t> active(n, 0);
	ok = ok && std::fwrite(active.data(), 1, n, f.get()) == n;
	if (!ok) {
		fail(CGP_ERR_INVALID, path + ": write failed");
	}
}

}  // namespace cgp
