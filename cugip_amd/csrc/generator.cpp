// Synthetic grid graph-cut instances: blob volume + Gaussian noise -> int32
// terminal and neighbour capacities. Integer-only (table lookups, SplitMix64
// counters, exact int64 ellipsoid tests), so a seed gives the same volume and
// the same capacities on every machine.
//
// The grid construction mirrors the reference's image -> graph step
// (benchmarks/graph_cut/graph_cut.cpp:32-98): one vertex per voxel, t-links from
// per-voxel likelihoods, and one edge per voxel towards each neighbour of the
// negative half-neighbourhood (VonNeumannNeighborhood<3> offsets 1-3,
// neighborhood.hpp:51-91; MooreNeighborhood<3> offsets 1-13, neighborhood.hpp:207-245).
// Voxels are numbered linearly with x fastest (access_utils.hpp:77-91).
#include "cugip_hip.h"
#include "gen_tables.h"
#include "grid_dirs.h"

#include <cstdint>
#include <cstdlib>
#include <cstring>

namespace {

inline uint64_t splitmix64(uint64_t x)
{
	x += 0x9E3779B97F4A7C15ull;
	x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
	x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
	return x ^ (x >> 31);
}

struct Blob {
	int64_t cx, cy, cz, a, b, c;
};

inline int noise_from_uniform(uint32_t u)
{
	// smallest j with u < CDF[j]; CDF is increasing and CDF[last] = 2^32
	int lo = 0, hi = 2 * CGP_NOISE_HALF;
	while (lo < hi) {
		int mid = (lo + hi) >> 1;
		if ((uint64_t)u < CGP_NOISE_CDF[mid]) {
			hi = mid;
		} else {
			lo = mid + 1;
		}
	}
	return lo - CGP_NOISE_HALF;
}

}  // namespace

extern "C" int
cgp_generate_volume(int nx, int ny, int nz, uint64_t seed, uint64_t instance, uint8_t *out)
{
	if (nx <= 0 || ny <= 0 || nz <= 0 || !out) {
		return CGP_ERR_INVALID;
	}
	const uint64_t key = splitmix64(splitmix64(seed) ^ (0x5851F42D4C957F2Dull * (instance + 1)));
	auto draw = [&](uint64_t i) { return splitmix64(key + i); };

	const int64_t m = nx < ny ? (nx < nz ? nx : nz) : (ny < nz ? ny : nz);
	int64_t amin = (8 * m) / 100;
	if (amin < 1) {
		amin = 1;
	}
	int64_t amax = (25 * m) / 100;
	if (amax < amin) {
		amax = amin;
	}
	const int count = 3 + (int)(draw(0) % 6);
	Blob blobs[8];
	for (int b = 0; b < count; ++b) {
		const uint64_t base = 1 + 7 * (uint64_t)b;
		blobs[b].cx = (int64_t)(draw(base + 0) % (uint64_t)nx);
		blobs[b].cy = (int64_t)(draw(base + 1) % (uint64_t)ny);
		blobs[b].cz = (int64_t)(draw(base + 2) % (uint64_t)nz);
		blobs[b].a = amin + (int64_t)(draw(base + 3) % (uint64_t)(amax - amin + 1));
		blobs[b].b = amin + (int64_t)(draw(base + 4) % (uint64_t)(amax - amin + 1));
		blobs[b].c = amin + (int64_t)(draw(base + 5) % (uint64_t)(amax - amin + 1));
	}
	const uint64_t noise_key = splitmix64(key ^ 0xD1B54A32D192ED03ull);
	for (int z = 0; z < nz; ++z) {
		for (int y = 0; y < ny; ++y) {
			for (int x = 0; x < nx; ++x) {
				int value = 60;
				for (int b = 0; b < count; ++b) {
					const Blob &e = blobs[b];
					const int64_t dx = x - e.cx, dy = y - e.cy, dz = z - e.cz;
					const int64_t a2 = e.a * e.a, b2 = e.b * e.b, c2 = e.c * e.c;
					// (dx/a)^2 + (dy/b)^2 + (dz/c)^2 <= 1, exactly, in int64
					if (dx * dx * b2 * c2 + dy * dy * a2 * c2 + dz * dz * a2 * b2 <= a2 * b2 * c2) {
						value = 180;
						break;
					}
				}
				const uint64_t idx = ((uint64_t)z * (uint64_t)ny + (uint64_t)y) * (uint64_t)nx + (uint64_t)x;
				const uint32_t u = (uint32_t)(splitmix64(noise_key + idx) >> 32);
				value += noise_from_uniform(u);
				value = value < 0 ? 0 : (value > 255 ? 255 : value);
				out[idx] = (uint8_t)value;
			}
		}
	}
	return CGP_OK;
}

extern "C" int
cgp_caps_from_volume(
	int nx, int ny, int nz, int conn,
	const uint8_t *volume,
	int32_t *cap_source,
	int32_t *cap_sink,
	int32_t *cap_n)
{
	if (nx <= 0 || ny <= 0 || nz <= 0 || !volume || !cap_source || !cap_sink || !cap_n) {
		return CGP_ERR_INVALID;
	}
	if (conn != 6 && conn != 26) {
		return CGP_ERR_INVALID;
	}
	const int half = conn / 2;
	const int8_t (*offs)[3] = conn == 6 ? CGP_HALF_OFFSETS_6 : CGP_HALF_OFFSETS_26;
	const size_t vcount = (size_t)nx * ny * nz;
	for (int z = 0; z < nz; ++z) {
		for (int y = 0; y < ny; ++y) {
			for (int x = 0; x < nx; ++x) {
				const size_t idx = ((size_t)z * ny + y) * nx + x;
				const int iv = volume[idx];
				cap_source[idx] = CGP_TLINK_SOURCE[iv];
				cap_sink[idx] = CGP_TLINK_SINK[iv];
				for (int j = 0; j < half; ++j) {
					const int ux = x + offs[j][0], uy = y + offs[j][1], uz = z + offs[j][2];
					int32_t w = 0;
					if (ux >= 0 && ux < nx && uy >= 0 && uy < ny && uz >= 0 && uz < nz) {
						const size_t uidx = ((size_t)uz * ny + uy) * nx + ux;
						const int diff = std::abs(iv - (int)volume[uidx]);
						const int cls = (offs[j][0] != 0) + (offs[j][1] != 0) + (offs[j][2] != 0) - 1;
						w = CGP_NLINK_WEIGHT[cls][diff];
					}
					cap_n[(size_t)j * vcount + idx] = w;
				}
			}
		}
	}
	return CGP_OK;
}

namespace cgp {

// The lookup tables of the synthetic instances as front-end weights: no markers, no border.
void default_volume_weights(cgp_volume_weights &w)
{
	std::memset(&w, 0, sizeof(w));
	for (int i = 0; i < 256; ++i) {
		w.tlink_source[i] = CGP_TLINK_SOURCE[i];
		w.tlink_sink[i] = CGP_TLINK_SINK[i];
		for (int c = 0; c < 3; ++c) {
			w.nlink[c][i] = CGP_NLINK_WEIGHT[c][i];
		}
	}
	w.marker_source = -1;
	w.marker_sink = -1;
	w.marker_capacity = 0;
	w.border = 0;
}

}  // namespace cgp
