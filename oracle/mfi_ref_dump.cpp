// Reads an MFI file with the reference's own reader (benchmarks/mfi.cpp, compiled from the
// reference sources by oracle/Makefile into oracle/_ref/) and prints its header fields and
// per-plane checksums, so tests can check the files this library writes against that reader.
// Output: one line "dimension connectivity width height depth maxflow symmetric active_count
// type_terminal type_neighbor", then one line per plane (source, sink, arcs..., active mask):
// "<sum> <sum of (i + 1) * value mod 2^64>".
#include <cstdint>
#include <cstdio>

#include "mfi.h"

static void plane(const void *data, int type, long n)
{
	unsigned long long sum = 0, weighted = 0;
	for (long i = 0; i < n; ++i) {
		unsigned long long v = type == 0 ? ((const uint8_t *)data)[i]
		                                 : (type == 1 ? ((const uint16_t *)data)[i] : ((const uint32_t *)data)[i]);
		sum += v;
		weighted += (unsigned long long)(i + 1) * v;
	}
	std::printf("%llu %llu\n", sum, weighted);
}

int main(int argc, char **argv)
{
	if (argc < 2) {
		return 2;
	}
	MFI *m = mfi_read(argv[1]);
	if (!m) {
		std::printf("unreadable\n");
		return 1;
	}
	const long n = (long)m->width * m->height * m->depth;
	std::printf("%d %d %d %d %d %ld %d %d %d %d\n", m->dimension, m->connectivity, m->width, m->height, m->depth,
	            m->maxflow, m->symmetric_arcs ? 1 : 0, m->active_count, (int)m->type_terminal_cap,
	            (int)m->type_neighbor_cap);
	plane(m->cap_source, m->type_terminal_cap, n);
	plane(m->cap_sink, m->type_terminal_cap, n);
	for (int a = 0; a < m->connectivity; ++a) {
		plane(m->cap_neighbor[a], m->type_neighbor_cap, n);
	}
	plane(m->active_mask, 0, n);
	mfi_free(m);
	return 0;
}
