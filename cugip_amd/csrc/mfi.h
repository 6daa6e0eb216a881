// MFI max-flow instance files: the regular-grid benchmark format the reference reads in
// benchmarks/mfi.h:4-77 and benchmarks/mfi.cpp:105-200. Host code only.
#pragma once

#include <cstdint>
#include <string>
#include <vector>

#include "cugip_hip.h"

namespace cgp {

struct MfiFile {
	cgp_mfi_info info;
	std::vector<uint32_t> source, sink;        // nodes
	std::vector<std::vector<uint32_t>> arcs;   // connectivity x nodes: arcs[a][v] = capacity v -> v + offset(a)
	std::vector<uint8_t> active;               // nodes
};

// Grid form of an instance: the layout of cgp_set_grid with both directions of every edge.
struct MfiGrid {
	int nx = 0, ny = 0, nz = 0, conn = 0;
	std::vector<int32_t> cs, ct, cn, cnb;      // cn, cnb: conn / 2 planes
};

MfiFile mfi_read(const std::string &path);
MfiGrid mfi_to_grid(const MfiFile &file);
// Writes a 3D instance (6- or 26-connected) with uint32 capacities; cnb may be null (symmetric).
void mfi_write_grid(const std::string &path, int nx, int ny, int nz, int conn, const int32_t *cs, const int32_t *ct,
                    const int32_t *cn, const int32_t *cnb, int64_t maxflow, const char *metadata);

}  // namespace cgp
