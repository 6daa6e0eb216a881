/*
 * cugip_hip.h -- C ABI of the MI355X (gfx950) push-relabel graph cut.
 *
 * This is the boundary between host code and the HIP implementation. The
 * header-only C++ drop-in (include/cugip/graph_cut.hpp, cugip::Graph<TFlow>)
 * forwards every call to these functions; any other host language binds them
 * directly (plain pointers and sizes, no C++ or torch types).
 *
 * Semantics (all capacities are non-negative int32, flows are exact int64):
 *   - A graph has n vertices plus the terminals s and t. Each vertex v has a
 *     source t-link s->v and a sink t-link v->t. Each undirected edge record
 *     (first, second), first < second, carries a capacity first->second (fwd)
 *     and second->first (bwd).
 *   - cgp_max_flow computes the maximum s-t flow value exactly.
 *   - cgp_fill_segments writes the reference's segmentation: `mask` for every
 *     vertex that can still reach t in the final residual graph (T_min, the
 *     vertices whose last BFS label is finite), `background` elsewhere.
 *   - cgp_source_set writes the canonical minimum cut: 1 for every vertex
 *     reachable from s in the residual graph of a maximum flow (S_min), else 0.
 *     S_min and T_min do not depend on which maximum flow was found.
 *
 * Every function returns CGP_OK (0) or a CGP_ERR_* code; cgp_last_error gives
 * the message. One handle is used by one host thread at a time; different
 * handles are independent (each owns a HIP stream and its device buffers).
 */
#ifndef CUGIP_HIP_H
#define CUGIP_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define CGP_OK 0
#define CGP_ERR_INVALID 1       /* bad argument, dimension or call order */
#define CGP_ERR_NOMEM 2         /* device or host allocation failed */
#define CGP_ERR_NOT_CONVERGED 3 /* round cap reached; no result is returned */
#define CGP_ERR_OVERFLOW 4      /* a capacity sum does not fit the int32 state */
#define CGP_ERR_DEVICE 5        /* HIP runtime error */
#define CGP_ERR_STATE 6         /* result requested before cgp_max_flow */

typedef struct cgp_graph_s *cgp_graph;

/* Layout-compatible with cugip::EdgeRecord (advanced_operations/detail/edge_record.hpp:10-29):
   `second` at byte offset 0, `first` at byte offset 4, first = min(endpoints). */
typedef struct cgp_edge {
	uint32_t second;
	uint32_t first;
} cgp_edge;

/* Run-time solver policy. Zero fields take the library default.
   Counterpart of cugip::GraphCutPolicy (advanced_operations/detail/graph_cut_policies.hpp:23-89). */
typedef struct cgp_policy {
	int32_t tile_x, tile_y, tile_z;       /* grid path: tile shape (only supported shapes are accepted) */
	int32_t discharges_per_relabel;       /* grid path: discharge phases between two global relabels */
	int32_t max_tile_iterations;          /* grid path: region relabel+push rounds per tile visit */
	int32_t max_rounds;                   /* global relabels before CGP_ERR_NOT_CONVERGED */
	float minimal_residual;               /* EdgeReverseTraversable::minimalResidual (edge_record.hpp:60-72);
	                                         > 0 closes thin edges during the relabel rounds, and the solve
	                                         then finishes with exact (threshold 0) rounds, so the result
	                                         stays exact */
	int32_t engine;                       /* grid path: CGP_ENGINE_* (0 = automatic) */
	int32_t reserved[7];
} cgp_policy;

/* Grid engines: the same tile kernels, driven two ways.
   TILES: the host runs the relabel / discharge / closure loop and reads tile counts between stages
   (MinCut::compute's host-driven loop, advanced_operations/detail/graph_cut_implementation.hpp:122-198).
   GRAPH: device-driven. A control block in device memory holds the solve's stage, every kernel is
   gated by it and controller kernels move it on; init plus one round of stages is captured in a HIP
   graph once and replayed, with one read of the control block per replay, until it says done.
   GRAPH serves 6-connected grids only and refuses trace hooks, profiling, minimal_residual > 0
   and a certificate (CGP_ERR_INVALID).
   AUTO: TILES (measured as fast or faster than GRAPH on every timed workload). */
#define CGP_ENGINE_AUTO 0
#define CGP_ENGINE_TILES 1
#define CGP_ENGINE_GRAPH 2

/* Per-solve counters and device timings. */
typedef struct cgp_stats {
	int64_t flow;              /* max-flow value */
	int32_t rounds;            /* global relabels (BFS from the sink) */
	int32_t bfs_passes;        /* tile passes (grid) or BFS levels (generic) over all rounds */
	int32_t discharge_phases;  /* grid: colour phases; generic: push levels */
	int32_t cut_passes;        /* passes of the S_min closure */
	int64_t bfs_tile_visits;   /* grid: tiles processed by the global relabel */
	int64_t discharge_tile_visits;
	int64_t cut_tile_visits;
	/* host wall time per phase; kernels run asynchronously, so a phase's time includes waiting for
	   the device work queued before it (e.g. the init kernel's time shows in relabel_ms, the
	   discharge phases' in cut_ms) */
	double init_ms, relabel_ms, discharge_ms, cut_ms, total_ms;
	int64_t voxels;            /* vertices (grid: voxels) */
	int32_t tile_x, tile_y, tile_z, conn;
	int32_t graph_replays;     /* GRAPH engine: replays of the round graph (0 on TILES) */
	int32_t reserved[7];
} cgp_stats;

/* Trace hooks of the reference's trace-object concept (advanced_operations/graph_cut.hpp:26-45). */
#define CGP_TRACE_STARTED 0
#define CGP_TRACE_BEGIN_ITERATION 1
#define CGP_TRACE_AFTER_RELABEL 2
#define CGP_TRACE_AFTER_PUSH 3
#define CGP_TRACE_FINISHED 4
/* level_starts: cumulative vertex counts per BFS level (levelStarts of the reference), valid
   for CGP_TRACE_AFTER_RELABEL only; flow: flow into t so far. */
typedef void (*cgp_trace_fn)(void *user, int32_t event, int32_t iteration, int32_t done, int64_t flow,
                             const int32_t *level_starts, int32_t level_count);

/* ---- handle ------------------------------------------------------------------------- */
int cgp_create(cgp_graph *out);
int cgp_destroy(cgp_graph graph);
/* Message of the last failed call on `graph`; with NULL, of the last failed call made without a
   handle (cgp_create, the MFI file functions). */
const char *cgp_last_error(cgp_graph graph);

/* ---- generic graphs: cugip::Graph<TFlow> (advanced_operations/graph_cut.hpp:47-146) ---- */
/* replaces Graph<TFlow>::set_vertex_count, graph_cut.hpp:213-237 */
int cgp_set_vertex_count(cgp_graph graph, int32_t vertex_count);
/* replaces Graph<TFlow>::set_nweights, graph_cut.hpp:239-290 (fwd = first->second, bwd = second->first) */
int cgp_set_nweights(cgp_graph graph, int32_t edge_count, const cgp_edge *edges, const int32_t *fwd,
                     const int32_t *bwd);
/* replaces Graph<TFlow>::set_tweights, graph_cut.hpp:292-302 */
int cgp_set_tweights(cgp_graph graph, const int32_t *source, const int32_t *sink);

/* ---- 3D grid graphs (implicit edges, no CSR) --------------------------------------------
   Voxel v = (z*ny + y)*nx + x. conn is 6 or 26. cap_n holds conn/2 planes of nx*ny*nz int32:
   plane j is the capacity v -> v + o_j where o_j is the j-th offset of the negative half
   neighbourhood (VonNeumannNeighborhood<3>::offset(1+j) for 6, MooreNeighborhood<3>::offset(1+j)
   for 26; cugip/neighborhood.hpp). cap_n_back (may be NULL = symmetric) holds v + o_j -> v.
   Entries whose neighbour falls outside the volume are ignored. This is the grid counterpart of
   the caller-side construction in benchmarks/graph_cut/graph_cut.cpp:32-98. */
int cgp_set_grid(cgp_graph graph, int32_t nx, int32_t ny, int32_t nz, int32_t conn, const int32_t *cap_source,
                 const int32_t *cap_sink, const int32_t *cap_n, const int32_t *cap_n_back);
/* Same, with all pointers in device memory of the handle's device (copied device to device). */
int cgp_set_grid_device(cgp_graph graph, int32_t nx, int32_t ny, int32_t nz, int32_t conn,
                        const int32_t *cap_source, const int32_t *cap_sink, const int32_t *cap_n,
                        const int32_t *cap_n_back);

/* ---- batches of independent 3D grids -------------------------------------------------------
   One handle holds `count` grids of one connectivity (each with the layout of cgp_set_grid) and
   solves them together: the tiles of every grid share the tile lists of each phase, so a launch
   serves all grids and small or uneven grids keep the device full. No counterpart in the
   reference (its Graph<TFlow> holds one graph); the batched entry of the grid path. After
   cgp_set_grid_batch, cgp_max_flow returns the total flow of all grids; cgp_fill_segments,
   cgp_source_set and cgp_labels are refused unless count is 1. */
typedef struct cgp_grid_instance {
	int32_t nx, ny, nz;
	int32_t reserved0;           /* 0 */
	const int32_t *cap_source;
	const int32_t *cap_sink;
	const int32_t *cap_n;
	const int32_t *cap_n_back;   /* NULL = symmetric */
} cgp_grid_instance;
/* device_pointers != 0: every capacity pointer is device memory of the handle's device. */
int cgp_set_grid_batch(cgp_graph graph, int32_t count, int32_t conn, const cgp_grid_instance *grids,
                       int32_t device_pointers);
int cgp_batch_size(cgp_graph graph, int32_t *count);
/* Solves every grid; flows (may be NULL) receives `count` exact values; stats cover the batch. */
int cgp_solve_batch(cgp_graph graph, int64_t *flows, cgp_stats *stats);
/* S_min and T_min (fill_segments semantics) of grid `index`, host buffers of its voxel count. */
int cgp_batch_source_set(cgp_graph graph, int32_t index, uint8_t *out);
int cgp_batch_fill_segments(cgp_graph graph, int32_t index, uint8_t *out, uint8_t mask, uint8_t background);

/* ---- solving ----------------------------------------------------------------------------- */
int cgp_set_policy(cgp_graph graph, const cgp_policy *policy);   /* NULL restores the defaults */
int cgp_get_policy(cgp_graph graph, cgp_policy *policy);
int cgp_set_trace(cgp_graph graph, cgp_trace_fn fn, void *user); /* fn NULL disables tracing */
/* replaces Graph<TFlow>::max_flow / max_flow_with_tracing, graph_cut.hpp:69-100 */
int cgp_max_flow(cgp_graph graph, int64_t *flow, cgp_stats *stats);
/* replaces Graph<TFlow>::fill_segments, graph_cut.hpp:318-333 (T_min; host buffer of n bytes) */
int cgp_fill_segments(cgp_graph graph, uint8_t *out, uint8_t mask, uint8_t background);
/* canonical minimum cut S_min as 0/1 bytes (host buffer of n bytes) */
int cgp_source_set(cgp_graph graph, uint8_t *out);
/* final vertex labels (exact BFS distance to t in the final residual graph, or
   CGP_LABEL_UNREACHABLE); debug view of Graph<TFlow>::mLabels, graph_cut.hpp:135. A grid solve
   ends by testing the S_min closure, not with a global relabel, so the first call after a solve
   (like cgp_fill_segments) runs one exact relabel. */
#define CGP_LABEL_UNREACHABLE 0x3fffffff
int cgp_labels(cgp_graph graph, int32_t *out);

/* ---- diagnostics and export -----------------------------------------------------------------
   What the reference's benchmark trace object reads from GraphCutData during and after a solve
   (benchmarks/graph_cut/graph_cut.cu:82-177: saturated edges, excess, labels), and the GraphML
   dump of Graph::save_to_graphml (advanced_operations/graph_to_graphml.hpp:6-61). */
/* State of grid `index` (0 for a single grid) once a solve has started, e.g. from a trace hook:
   excess (int32), labels (int32, CGP_LABEL_UNREACHABLE when t is unreachable; during a solve the
   labels as they stand, after it the exact labels of the final residual graph) and saturated directions (uint32 bit d set when the edge to the
   neighbour in residual direction d has capacity > 0 and residual 0; d < conn/2 are the offsets of
   cap_n, d >= conn/2 their mirrors). Any output may be NULL. */
int cgp_grid_state(cgp_graph graph, int32_t index, int32_t *excess, int32_t *labels, uint32_t *saturated);
/* Residual capacities of grid `index` (0 for a single grid) once a solve has started: the residuals
   of the reference's GraphCutData (advanced_operations/detail/graph_cut_data.hpp:20-140) in the grid
   layout. residuals receives conn planes of voxel-count int32, plane d = residual v -> v + o_d
   (d < conn/2: the offsets of cap_n; d >= conn/2: their mirrors; 0 where the neighbour is outside);
   sink_residual the residual v -> t. Either may be NULL. */
int cgp_grid_residuals(cgp_graph graph, int32_t index, int32_t *residuals, int32_t *sink_residual);
/* Capacity of the cut whose source side is {s} + {v : source_side[v] != 0} of grid `index`
   (host buffer of its voxel count): the sum of cs over voxels outside, ct over voxels inside,
   and the capacities of the n-links from inside to outside. Exact int64. A cut whose capacity
   equals the max flow is a minimum cut. */
int cgp_cut_capacity(cgp_graph graph, int32_t index, const uint8_t *source_side, int64_t *capacity);
/* Writes the graph (generic or single grid) as GraphML: nodes "source", "sink", "n<i>"; one
   directed edge per positive capacity (both directions of every n-link, source -> n<i>,
   n<i> -> sink), capacity in the "capacity" key. */
int cgp_save_graphml(cgp_graph graph, const char *path);

/* ---- profiling -----------------------------------------------------------------------------
   When enabled, the grid solver brackets every kernel launch with HIP events on its own stream
   and waits for each launch, so per-kernel device time and algorithmic bytes can be read back
   after cgp_max_flow. This serialises the solve: leave it off when timing whole solves.
   algorithmic_bytes = voxel_visits x bytes_per_voxel, where bytes_per_voxel counts each plane
   a kernel must read or write once per voxel it processes (halo values staged in LDS count 0). */
typedef struct cgp_kernel_profile {
	char name[32];
	int32_t launches;
	int32_t reserved0;
	int64_t voxel_visits;
	double device_ms;
	double algorithmic_bytes;
	double bytes_per_voxel;
} cgp_kernel_profile;
int cgp_set_profiling(cgp_graph graph, int32_t enable);
int cgp_get_profile(cgp_graph graph, cgp_kernel_profile *out, int32_t capacity, int32_t *count);

/* ---- MFI instance files --------------------------------------------------------------------
   The regular-grid max-flow benchmark format the reference reads (benchmarks/mfi.h:4-77,
   benchmarks/mfi.cpp:105-200): 2D 4/8-connected and 3D 6/18/26-connected grids with uint8, uint16
   or uint32 capacities and the instance's known max flow in the header. An instance loads as a
   6-connected grid (2D-4C, 3D-6C; depth 1 for 2D) or a 26-connected grid (2D-8C, 3D-18C, 3D-26C;
   the arcs a lower connectivity lacks have capacity 0), with both directions of every edge. */
typedef struct cgp_mfi_info {
	int32_t dimension, connectivity;       /* as stored */
	int32_t width, height, depth;          /* nx, ny, nz of the grid */
	int32_t symmetric_arcs;
	int64_t maxflow;                       /* header value */
	int32_t active_count;
	int32_t type_terminal_cap, type_neighbor_cap; /* 0 uint8, 1 uint16, 2 uint32 */
	int32_t grid_conn;                     /* 6 or 26 */
	int32_t reserved[4];
	char metadata[4096];
} cgp_mfi_info;
/* Header of an MFI file. Without a handle: on failure cgp_last_error(NULL) has the message. */
int cgp_mfi_read_info(const char *path, cgp_mfi_info *info);
/* Capacities of an MFI file in the layout of cgp_set_grid: host buffers of width*height*depth
   int32 (cap_n and cap_n_back: grid_conn/2 such planes). */
int cgp_mfi_to_grid(const char *path, int32_t *cap_source, int32_t *cap_sink, int32_t *cap_n, int32_t *cap_n_back);
/* Reads an MFI file into the handle (as cgp_set_grid); info (may be NULL) receives its header. */
int cgp_set_grid_mfi(cgp_graph graph, const char *path, cgp_mfi_info *info);
/* Writes a 3D grid (layout of cgp_set_grid, cap_n_back NULL = symmetric) as an MFI file with
   uint32 capacities, `maxflow` in the header and optional metadata text. */
int cgp_write_mfi(const char *path, int32_t nx, int32_t ny, int32_t nz, int32_t conn, const int32_t *cap_source,
                  const int32_t *cap_sink, const int32_t *cap_n, const int32_t *cap_n_back, int64_t maxflow,
                  const char *metadata);

/* ---- region adjacency graphs ------------------------------------------------------------------
   The text format of the reference's region benchmarks (benchmarks/graph_cut/graph.hpp:48-104,
   loadGraph / saveGraph; marker files of benchmarks/graph_cut/segmentation_regions.cpp:53-80):
   graph lines "n;<id>;<count>;<sum>" (regions, ids 1..n) and "e;<a>;<b>;<count>;<sum>" (adjacent
   regions and their boundary statistic), marker lines "f <id>" (object) and "b <id>" (background);
   '#' lines and empty lines are skipped, and a later edge record of the same (a, b) replaces the
   earlier one. As computeCudaGraphCut builds its GraphData (benchmarks/graph_cut/graph_cut.cpp:
   100-154), region id i is vertex i - 1, every edge has weight round(sum * weight_scale) in both
   directions, background markers get marker_capacity on the source link and object markers on the
   sink link; the object segmentation is then cgp_fill_segments (the vertices that still reach t). */
int cgp_region_graph_info(const char *graph_path, int32_t *vertex_count, int32_t *edge_count);
/* Fills edges and weights (edge_count entries) and the t-links (vertex_count entries) of a graph
   file and a marker file (markers_path may be NULL: no markers), ready for cgp_set_vertex_count,
   cgp_set_nweights(edges, weights, weights) and cgp_set_tweights. Host buffers. */
int cgp_region_graph_read(const char *graph_path, const char *markers_path, double weight_scale,
                          int32_t marker_capacity, cgp_edge *edges, int32_t *weights, int32_t *cap_source,
                          int32_t *cap_sink);

/* ---- on-device image -> graph front end -----------------------------------------------------
   Capacities computed on the device from a uint8 volume, so that they never exist on the host:
   the counterpart of the caller-side construction in benchmarks/graph_cut/graph_cut.cpp:32-98
   (t-links from markers, border voxels forced to the source, contrast-sensitive n-links).
   Every capacity is a table lookup: t-links by intensity, n-links by |I_v - I_u| per neighbour
   class (0 face, 1 edge, 2 corner neighbour). Then, in this order: a voxel whose marker equals
   marker_source gets t-links (marker_capacity, 0) and one equal to marker_sink (0,
   marker_capacity); a voxel with some coordinate c < border or c >= size - border gets
   (marker_capacity, 0). The reference's scheme is marker_source 128, marker_sink 255,
   border 2, zero t-link tables. */
typedef struct cgp_volume_weights {
	int32_t tlink_source[256];
	int32_t tlink_sink[256];
	int32_t nlink[3][256];
	int32_t marker_source, marker_sink;   /* -1: no such marker */
	int32_t marker_capacity;
	int32_t border;                       /* 0: no border forcing */
	int32_t reserved[4];
} cgp_volume_weights;
/* The tables of the synthetic instances (those of cgp_caps_from_volume), no markers, no border. */
int cgp_default_volume_weights(cgp_volume_weights *weights);
/* Sets a grid from a volume (and optional marker volume, NULL = none) of nx*ny*nz bytes; weights
   NULL = the defaults. device_pointers != 0: volume and markers are device memory. */
int cgp_set_grid_volume(cgp_graph graph, int32_t nx, int32_t ny, int32_t nz, int32_t conn, const uint8_t *volume,
                        const uint8_t *markers, const cgp_volume_weights *weights, int32_t device_pointers);
/* The input capacities the handle holds for grid `index` (0 for a single grid), in the layout of
   cgp_set_grid; cap_n_back receives the forward planes when the grid is symmetric. */
int cgp_grid_capacities(cgp_graph graph, int32_t index, int32_t *cap_source, int32_t *cap_sink, int32_t *cap_n,
                        int32_t *cap_n_back);

/* ---- learned cut proposal (3D CNN on the matrix cores) --------------------------------------
   Four 3x3x3 convolutions with 16 channels and ReLU, then a 1x1x1 sigmoid head; BF16 operands,
   FP32 accumulation (v_mfma_f32_16x16x16_bf16). Input per voxel: intensity / 255, source and sink
   t-link / 1000. Parameters (fp32, CGP_CNN_PARAMS): W1[27][4][16] (tap, input channel, output
   channel; input channel 3 unused), b1[16], then W2, W3, W4 [27][16][16] each followed by its bias
   [16], then w_head[16], b_head; tap t = (dz+1)*9 + (dy+1)*3 + (dx+1). The network reads grid 0 of
   a graph handle (its t-links, already on the device) plus the volume. The thresholded output is
   a candidate source side whose cut capacity bounds the max flow from above: a certificate.
   No counterpart in the reference. */
#define CGP_CNN_PARAMS 22545
typedef struct cgp_cnn_s *cgp_cnn;
int cgp_cnn_create(cgp_cnn *out);
int cgp_cnn_destroy(cgp_cnn cnn);
const char *cgp_cnn_last_error(cgp_cnn cnn);
int cgp_cnn_set_params(cgp_cnn cnn, const float *params, int32_t count);
int cgp_cnn_get_params(cgp_cnn cnn, float *params, int32_t count);
/* Prediction for the graph's grid from its volume (host bytes; NULL = the volume the grid was built
   from by cgp_set_grid_volume, already on the device). side (host, 0/1 bytes) and prob (host
   floats) may be NULL; certificate (may be NULL) receives the exact capacity of the cut. The same
   volume rule holds for cgp_cnn_gradients and cgp_cnn_train_step. */
int cgp_cnn_predict(cgp_cnn cnn, cgp_graph graph, const uint8_t *volume, uint8_t *side, float *prob,
                    int64_t *certificate);
/* Mean binary cross-entropy against the graph's S_min (the graph must be solved) and its gradient
   (grad: CGP_CNN_PARAMS floats, may be NULL). */
int cgp_cnn_gradients(cgp_cnn cnn, cgp_graph graph, const uint8_t *volume, double *loss, float *grad);
/* One Adam step (beta1 0.9, beta2 0.999, eps 1e-8) on that gradient; loss before the step. */
int cgp_cnn_train_step(cgp_cnn cnn, cgp_graph graph, const uint8_t *volume, float learning_rate, double *loss);
/* Solve with a certificate (the capacity of some cut): stops as soon as the flow into t reaches
   it, runs to completion otherwise; stopped_early (may be NULL) says which. Results are exact
   either way; fill_segments after an early stop runs one exact relabel first. */
int cgp_max_flow_certified(cgp_graph graph, int64_t certificate, int64_t *flow, cgp_stats *stats,
                           int32_t *stopped_early);

/* ---- synthetic instances ------------------------------------------------------------------
   Blob volume (3-8 ellipsoids at 180 on 60, N(0, 25^2) noise, clamped to 0-255) and the int32
   capacities derived from it (Boykov-Jolly t-links, contrast-sensitive n-links). Integer-only:
   the same arguments give the same bytes everywhere. */
int cgp_generate_volume(int32_t nx, int32_t ny, int32_t nz, uint64_t seed, uint64_t instance, uint8_t *out);
int cgp_caps_from_volume(int32_t nx, int32_t ny, int32_t nz, int32_t conn, const uint8_t *volume,
                         int32_t *cap_source, int32_t *cap_sink, int32_t *cap_n);

/* ---- solve without allocating ---------------------------------------------------------------
   No counterpart in the reference; for callers that cannot allocate device memory (tight serving
   loops, harnesses that forbid hipMalloc). One grid solve whose working memory (per-voxel residual
   state, tile flags and lists, counters) is a fixed scratch array of CGP_NOALLOC_SCRATCH_MIB MiB
   that each GPU the library runs on holds from its first such call until the process ends; the
   call itself makes no device or pinned allocation. Capacities: device arrays in the layout of
   cgp_set_grid, read in place (not copied). source_set (device memory of nx*ny*nz bytes, may be
   NULL) receives S_min as 0/1 bytes, flow the exact max-flow value. stream: a hipStream_t (NULL =
   the default stream); the call returns when its work on it is done. A grid whose state does not
   fit the scratch (41 bytes per voxel 6-connected, 121 bytes 26-connected, plus a few bytes per
   tile: about 1.6 M or 0.55 M voxels) fails with CGP_ERR_INVALID. Every call on one GPU uses that
   GPU's one scratch array, which keeps the last call's contents (the solve clears what it needs);
   the library runs such calls one at a time, so calls from several host threads queue. Errors:
   cgp_last_error(NULL). */
#define CGP_NOALLOC_SCRATCH_MIB 64
int cgp_grid_max_flow_noalloc(int32_t nx, int32_t ny, int32_t nz, int32_t conn, const int32_t *cap_source,
                              const int32_t *cap_sink, const int32_t *cap_n, const int32_t *cap_n_back,
                              uint8_t *source_set, int64_t *flow, void *stream);

/* ---- device ------------------------------------------------------------------------------- */
int cgp_device_count(int32_t *count);
/* Version of this ABI; bumps on any incompatible change. */
int32_t cgp_abi_version(void);

#ifdef __cplusplus
}
#endif

#endif /* CUGIP_HIP_H */
