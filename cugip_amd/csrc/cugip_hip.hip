// C ABI of the graph cut (include/cugip_hip.h): handle management, argument
// checks, error capture, and dispatch to the grid or the generic (CSR) solver.
#include <cstdio>
#include <cstring>
#include <memory>
#include <mutex>
#include <string>
#include <vector>

#include "common.h"
#include "cnn.h"
#include "grid_dirs.h"
#include "mfi.h"
#include "solver.h"

struct cgp_graph_s {
	enum class Mode { None, Generic, Grid };
	std::string last_error;
	hipStream_t stream = nullptr;
	int device = 0;
	Mode mode = Mode::None;
	cgp_policy policy = cgp::default_policy();
	bool profiling = false;
	cgp::TraceHook trace;
	std::unique_ptr<cgp::CsrSolver> csr;
	std::unique_ptr<cgp::GridSolver> grid;
};

struct cgp_cnn_s {
	std::string last_error;
	hipStream_t stream = nullptr;
	int device = 0;
	std::unique_ptr<cgp::CnnModel> model;
};

namespace {

std::mutex g_create_error_mutex;
std::string g_create_error;

constexpr int32_t kAbiVersion = 2;

template <typename Fn>
int guarded(cgp_graph g, Fn &&fn)
{
	if (!g) {
		return CGP_ERR_INVALID;
	}
	try {
		CGP_HIP(hipSetDevice(g->device));
		fn();
		g->last_error.clear();
		return CGP_OK;
	} catch (const cgp::Error &e) {
		g->last_error = e.what();
		return e.code;
	} catch (const std::bad_alloc &) {
		g->last_error = "host allocation failed";
		return CGP_ERR_NOMEM;
	} catch (const std::exception &e) {
		g->last_error = e.what();
		return CGP_ERR_DEVICE;
	}
}

// Calls without a handle report their error through cgp_last_error(NULL).
template <typename Fn>
int guarded_global(Fn &&fn)
{
	std::string message;
	int code = CGP_OK;
	try {
		fn();
		return CGP_OK;
	} catch (const cgp::Error &e) {
		message = e.what();
		code = e.code;
	} catch (const std::bad_alloc &) {
		message = "host allocation failed";
		code = CGP_ERR_NOMEM;
	} catch (const std::exception &e) {
		message = e.what();
		code = CGP_ERR_DEVICE;
	}
	std::lock_guard<std::mutex> lock(g_create_error_mutex);
	g_create_error = message;
	return code;
}

cgp::GridSolver &grid_of(cgp_graph g)
{
	if (g->mode != cgp_graph_s::Mode::Grid || !g->grid) {
		cgp::fail(CGP_ERR_STATE, "no grid set: call cgp_set_grid or cgp_set_grid_batch first");
	}
	return *g->grid;
}

// The grid of a single-grid handle; a batch must name the grid it means.
cgp::GridSolver &single_grid(cgp_graph g)
{
	cgp::GridSolver &grid = grid_of(g);
	if (grid.instance_count() != 1) {
		cgp::fail(CGP_ERR_INVALID, "the handle holds a batch of " + std::to_string(grid.instance_count()) +
		                               " grids: use the cgp_batch_* accessors");
	}
	return grid;
}

template <typename Fn>
int cnn_guarded(cgp_cnn c, Fn &&fn)
{
	if (!c) {
		return CGP_ERR_INVALID;
	}
	try {
		CGP_HIP(hipSetDevice(c->device));
		fn();
		c->last_error.clear();
		return CGP_OK;
	} catch (const cgp::Error &e) {
		c->last_error = e.what();
		return e.code;
	} catch (const std::bad_alloc &) {
		c->last_error = "host allocation failed";
		return CGP_ERR_NOMEM;
	} catch (const std::exception &e) {
		c->last_error = e.what();
		return CGP_ERR_DEVICE;
	}
}

cgp::GridSolver &cnn_grid(cgp_graph g)
{
	if (!g) {
		cgp::fail(CGP_ERR_INVALID, "null graph");
	}
	return single_grid(g);
}

cgp::CsrSolver &generic(cgp_graph g)
{
	if (g->mode != cgp_graph_s::Mode::Generic || !g->csr) {
		cgp::fail(CGP_ERR_STATE, "no generic graph: call cgp_set_vertex_count first");
	}
	return *g->csr;
}

}  // namespace

extern "C" {

int32_t cgp_abi_version(void) { return kAbiVersion; }

int cgp_device_count(int32_t *count)
{
	if (!count) {
		return CGP_ERR_INVALID;
	}
	int n = 0;
	const hipError_t err = hipGetDeviceCount(&n);
	if (err != hipSuccess) {
		(void)hipGetLastError();
		*count = 0;
		return CGP_ERR_DEVICE;
	}
	*count = n;
	return CGP_OK;
}

int cgp_create(cgp_graph *out)
{
	if (!out) {
		return CGP_ERR_INVALID;
	}
	*out = nullptr;
	std::unique_ptr<cgp_graph_s> g;
	try {
		g.reset(new cgp_graph_s());
		int n = 0;
		CGP_HIP(hipGetDeviceCount(&n));
		if (n <= 0) {
			cgp::fail(CGP_ERR_DEVICE, "no HIP device available");
		}
		CGP_HIP(hipGetDevice(&g->device));
		CGP_HIP(hipStreamCreateWithFlags(&g->stream, hipStreamNonBlocking));
	} catch (const cgp::Error &e) {
		std::lock_guard<std::mutex> lock(g_create_error_mutex);
		g_create_error = e.what();
		return e.code;
	} catch (const std::bad_alloc &) {
		return CGP_ERR_NOMEM;
	}
	*out = g.release();
	return CGP_OK;
}

int cgp_destroy(cgp_graph g)
{
	if (!g) {
		return CGP_OK;
	}
	(void)hipSetDevice(g->device);
	g->csr.reset();
	g->grid.reset();
	if (g->stream) {
		(void)hipStreamSynchronize(g->stream);
		(void)hipStreamDestroy(g->stream);
	}
	delete g;
	return CGP_OK;
}

const char *cgp_last_error(cgp_graph g)
{
	if (!g) {
		std::lock_guard<std::mutex> lock(g_create_error_mutex);
		return g_create_error.c_str();
	}
	return g->last_error.c_str();
}

int cgp_set_vertex_count(cgp_graph g, int32_t vertex_count)
{
	return guarded(g, [&] {
		if (vertex_count <= 0) {
			cgp::fail(CGP_ERR_INVALID, "vertex count must be positive");
		}
		g->grid.reset();
		if (!g->csr) {
			g->csr.reset(new cgp::CsrSolver(g->stream));
		}
		g->mode = cgp_graph_s::Mode::Generic;
		g->csr->set_vertex_count(vertex_count);
	});
}

int cgp_set_nweights(cgp_graph g, int32_t edge_count, const cgp_edge *edges, const int32_t *fwd, const int32_t *bwd)
{
	return guarded(g, [&] { generic(g).set_nweights(edge_count, edges, fwd, bwd); });
}

int cgp_set_tweights(cgp_graph g, const int32_t *source, const int32_t *sink)
{
	return guarded(g, [&] { generic(g).set_tweights(source, sink); });
}

static int set_grid_impl(cgp_graph g, int32_t nx, int32_t ny, int32_t nz, int32_t conn, const int32_t *cs,
                         const int32_t *ct, const int32_t *cn, const int32_t *cnb, bool device)
{
	return guarded(g, [&] {
		g->csr.reset();
		if (!g->grid) {
			g->grid.reset(new cgp::GridSolver(g->stream));
		}
		g->mode = cgp_graph_s::Mode::Grid;
		g->grid->set_grid(nx, ny, nz, conn, cs, ct, cn, cnb, device);
	});
}

int cgp_set_grid(cgp_graph g, int32_t nx, int32_t ny, int32_t nz, int32_t conn, const int32_t *cap_source,
                 const int32_t *cap_sink, const int32_t *cap_n, const int32_t *cap_n_back)
{
	return set_grid_impl(g, nx, ny, nz, conn, cap_source, cap_sink, cap_n, cap_n_back, false);
}

int cgp_set_grid_device(cgp_graph g, int32_t nx, int32_t ny, int32_t nz, int32_t conn, const int32_t *cap_source,
                        const int32_t *cap_sink, const int32_t *cap_n, const int32_t *cap_n_back)
{
	return set_grid_impl(g, nx, ny, nz, conn, cap_source, cap_sink, cap_n, cap_n_back, true);
}

int cgp_mfi_read_info(const char *path, cgp_mfi_info *info)
{
	return guarded_global([&] {
		if (!path || !info) {
			cgp::fail(CGP_ERR_INVALID, "null argument");
		}
		*info = cgp::mfi_read(path).info;
	});
}

int cgp_mfi_to_grid(const char *path, int32_t *cap_source, int32_t *cap_sink, int32_t *cap_n, int32_t *cap_n_back)
{
	return guarded_global([&] {
		if (!path || !cap_source || !cap_sink || !cap_n || !cap_n_back) {
			cgp::fail(CGP_ERR_INVALID, "null argument");
		}
		const cgp::MfiGrid grid = cgp::mfi_to_grid(cgp::mfi_read(path));
		std::memcpy(cap_source, grid.cs.data(), grid.cs.size() * sizeof(int32_t));
		std::memcpy(cap_sink, grid.ct.data(), grid.ct.size() * sizeof(int32_t));
		std::memcpy(cap_n, grid.cn.data(), grid.cn.size() * sizeof(int32_t));
		std::memcpy(cap_n_back, grid.cnb.data(), grid.cnb.size() * sizeof(int32_t));
	});
}

int cgp_set_grid_mfi(cgp_graph g, const char *path, cgp_mfi_info *info)
{
	return guarded(g, [&] {
		if (!path) {
			cgp::fail(CGP_ERR_INVALID, "null path");
		}
		const cgp::MfiFile file = cgp::mfi_read(path);
		const cgp::MfiGrid grid = cgp::mfi_to_grid(file);
		g->csr.reset();
		if (!g->grid) {
			g->grid.reset(new cgp::GridSolver(g->stream));
		}
		g->mode = cgp_graph_s::Mode::Grid;
		g->grid->set_grid(grid.nx, grid.ny, grid.nz, grid.conn, grid.cs.data(), grid.ct.data(), grid.cn.data(),
		                  grid.cnb.data(), false);
		if (info) {
			*info = file.info;
		}
	});
}

int cgp_write_mfi(const char *path, int32_t nx, int32_t ny, int32_t nz, int32_t conn, const int32_t *cap_source,
                  const int32_t *cap_sink, const int32_t *cap_n, const int32_t *cap_n_back, int64_t maxflow,
                  const char *metadata)
{
	return guarded_global([&] {
		if (!path) {
			cgp::fail(CGP_ERR_INVALID, "null path");
		}
		cgp::mfi_write_grid(path, nx, ny, nz, conn, cap_source, cap_sink, cap_n, cap_n_back, maxflow, metadata);
	});
}

int cgp_region_graph_info(const char *graph_path, int32_t *vertex_count, int32_t *edge_count)
{
	return guarded_global([&] {
		if (!graph_path || !vertex_count || !edge_count) {
			cgp::fail(CGP_ERR_INVALID, "null argument");
		}
		const cgp::RegionGraph g = cgp::read_region_graph(graph_path);
		*vertex_count = g.vertices;
		*edge_count = (int32_t)g.a.size();
	});
}

int cgp_region_graph_read(const char *graph_path, const char *markers_path, double weight_scale,
                          int32_t marker_capacity, cgp_edge *edges, int32_t *weights, int32_t *cap_source,
                          int32_t *cap_sink)
{
	return guarded_global([&] {
		if (!graph_path || !cap_source || !cap_sink) {
			cgp::fail(CGP_ERR_INVALID, "null argument");
		}
		const cgp::RegionGraph g = cgp::read_region_graph(graph_path);
		if (!g.a.empty() && (!edges || !weights)) {
			cgp::fail(CGP_ERR_INVALID, "null edge arrays");
		}
		std::vector<std::pair<bool, int32_t>> markers;
		if (markers_path) {
			markers = cgp::read_region_markers(markers_path);
		}
		cgp::region_graph_arrays(g, markers, weight_scale, marker_capacity, edges, weights, cap_source, cap_sink);
	});
}

int cgp_default_volume_weights(cgp_volume_weights *w)
{
	return guarded_global([&] {
		if (!w) {
			cgp::fail(CGP_ERR_INVALID, "null argument");
		}
		cgp::default_volume_weights(*w);
	});
}

int cgp_set_grid_volume(cgp_graph g, int32_t nx, int32_t ny, int32_t nz, int32_t conn, const uint8_t *volume,
                        const uint8_t *markers, const cgp_volume_weights *weights, int32_t device_pointers)
{
	return guarded(g, [&] {
		cgp_volume_weights defaults;
		if (!weights) {
			cgp::default_volume_weights(defaults);
			weights = &defaults;
		}
		g->csr.reset();
		if (!g->grid) {
			g->grid.reset(new cgp::GridSolver(g->stream));
		}
		g->mode = cgp_graph_s::Mode::Grid;
		g->grid->set_grid_volume(nx, ny, nz, conn, volume, markers, *weights, device_pointers != 0);
	});
}

int cgp_grid_capacities(cgp_graph g, int32_t index, int32_t *cap_source, int32_t *cap_sink, int32_t *cap_n,
                        int32_t *cap_n_back)
{
	return guarded(g, [&] {
		if (!cap_source || !cap_sink || !cap_n || !cap_n_back) {
			cgp::fail(CGP_ERR_INVALID, "null argument");
		}
		grid_of(g).capacities(index, cap_source, cap_sink, cap_n, cap_n_back);
	});
}

int cgp_grid_state(cgp_graph g, int32_t index, int32_t *excess, int32_t *labels, uint32_t *saturated)
{
	return guarded(g, [&] { grid_of(g).state(index, excess, labels, saturated); });
}

int cgp_grid_residuals(cgp_graph g, int32_t index, int32_t *residuals, int32_t *sink_residual)
{
	return guarded(g, [&] { grid_of(g).residuals(index, residuals, sink_residual); });
}

int cgp_cut_capacity(cgp_graph g, int32_t index, const uint8_t *source_side, int64_t *capacity)
{
	return guarded(g, [&] {
		if (!source_side || !capacity) {
			cgp::fail(CGP_ERR_INVALID, "null argument");
		}
		*capacity = grid_of(g).cut_capacity(index, source_side);
	});
}

int cgp_save_graphml(cgp_graph g, const char *path)
{
	return guarded(g, [&] {
		if (!path) {
			cgp::fail(CGP_ERR_INVALID, "null path");
		}
		std::vector<int32_t> tail, head, cap, cs, ct;
		if (g->mode == cgp_graph_s::Mode::Grid) {
			cgp::GridSolver &grid = single_grid(g);
			const cgp::InstanceDesc &d = grid.grid(0);
			const int conn = grid.conn(), half = conn / 2;
			const size_t n = (size_t)d.nvox;
			cs.resize(n);
			ct.resize(n);
			std::vector<int32_t> cn(n * half), cnb(n * half);
			grid.capacities(0, cs.data(), ct.data(), cn.data(), cnb.data());
			const int8_t(*offs)[3] = conn == 6 ? CGP_HALF_OFFSETS_6 : CGP_HALF_OFFSETS_26;
			for (int j = 0; j < half; ++j) {
				for (int z = 0; z < d.nz; ++z) {
					for (int y = 0; y < d.ny; ++y) {
						for (int x = 0; x < d.nx; ++x) {
							const int ux = x + offs[j][0], uy = y + offs[j][1], uz = z + offs[j][2];
							if (ux < 0 || ux >= d.nx || uy < 0 || uy >= d.ny || uz < 0 || uz >= d.nz) {
								continue;
							}
							const int32_t v = (int32_t)(((int64_t)z * d.ny + y) * d.nx + x);
							const int32_t u = (int32_t)(((int64_t)uz * d.ny + uy) * d.nx + ux);
							const int32_t f = cn[(size_t)j * n + (size_t)v], b = cnb[(size_t)j * n + (size_t)v];
							if (f > 0) {
								tail.push_back(v);
								head.push_back(u);
								cap.push_back(f);
							}
							if (b > 0) {
								tail.push_back(u);
								head.push_back(v);
								cap.push_back(b);
							}
						}
					}
				}
			}
		} else {
			generic(g).export_graph(tail, head, cap, cs, ct);
		}
		std::FILE *f = std::fopen(path, "w");
		if (!f) {
			cgp::fail(CGP_ERR_INVALID, std::string(path) + ": cannot create");
		}
		std::fprintf(f, "<?xml version=\"1.0\" encoding=\"UTF-8\"?>\n"
		                "<graphml xmlns=\"http://graphml.graphdrawing.org/xmlns\">\n"
		                "<key id=\"capacity\" for=\"edge\" attr.name=\"capacity\" attr.type=\"long\"/>\n"
		                "<graph id=\"G\" edgedefault=\"directed\">\n<node id=\"source\"/>\n<node id=\"sink\"/>\n");
		for (size_t i = 0; i < cs.size(); ++i) {
			std::fprintf(f, "<node id=\"n%zu\"/>\n", i);
		}
		const char *edge = "<edge source=\"%s%d\" target=\"%s%d\"><data key=\"capacity\">%d</data></edge>\n";
		for (size_t i = 0; i < tail.size(); ++i) {
			std::fprintf(f, edge, "n", tail[i], "n", head[i], cap[i]);
		}
		for (size_t i = 0; i < cs.size(); ++i) {
			if (cs[i] > 0) {
				std::fprintf(f, "<edge source=\"source\" target=\"n%zu\"><data key=\"capacity\">%d</data></edge>\n", i,
				             cs[i]);
			}
			if (ct[i] > 0) {
				std::fprintf(f, "<edge source=\"n%zu\" target=\"sink\"><data key=\"capacity\">%d</data></edge>\n", i,
				             ct[i]);
			}
		}
		const bool ok = std::fprintf(f, "</graph>\n</graphml>\n") > 0;
		if (std::fclose(f) != 0 || !ok) {
			cgp::fail(CGP_ERR_INVALID, std::string(path) + ": write failed");
		}
	});
}

int cgp_cnn_create(cgp_cnn *out)
{
	if (!out) {
		return CGP_ERR_INVALID;
	}
	*out = nullptr;
	return guarded_global([&] {
		std::unique_ptr<cgp_cnn_s> c(new cgp_cnn_s());
		int count = 0;
		CGP_HIP(hipGetDeviceCount(&count));
		if (count < 1) {
			cgp::fail(CGP_ERR_DEVICE, "no HIP device");
		}
		CGP_HIP(hipGetDevice(&c->device));
		CGP_HIP(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));
		c->model.reset(new cgp::CnnModel(c->stream));
		*out = c.release();
	});
}

int cgp_cnn_destroy(cgp_cnn c)
{
	if (!c) {
		return CGP_OK;
	}
	(void)hipSetDevice(c->device);
	c->model.reset();
	if (c->stream) {
		(void)hipStreamDestroy(c->stream);
	}
	delete c;
	return CGP_OK;
}

const char *cgp_cnn_last_error(cgp_cnn c)
{
	return c ? c->last_error.c_str() : "null handle";
}

int cgp_cnn_set_params(cgp_cnn c, const float *params, int32_t count)
{
	return cnn_guarded(c, [&] {
		if (!params || count != cgp::kCnnParams) {
			cgp::fail(CGP_ERR_INVALID, "expected " + std::to_string(cgp::kCnnParams) + " parameters");
		}
		c->model->set_params(params);
	});
}

int cgp_cnn_get_params(cgp_cnn c, float *params, int32_t count)
{
	return cnn_guarded(c, [&] {
		if (!params || count != cgp::kCnnParams) {
			cgp::fail(CGP_ERR_INVALID, "expected " + std::to_string(cgp::kCnnParams) + " parameters");
		}
		c->model->get_params(params);
	});
}

int cgp_cnn_predict(cgp_cnn c, cgp_graph g, const uint8_t *volume, uint8_t *side, float *prob, int64_t *certificate)
{
	return cnn_guarded(c, [&] {
		cgp::GridSolver &grid = cnn_grid(g);
		c->model->forward(grid, volume, false, false);
		c->model->prediction_to_host(side, prob);
		if (certificate) {
			*certificate = grid.cut_capacity_device(0, c->model->predicted_side());
		}
	});
}

int cgp_cnn_gradients(cgp_cnn c, cgp_graph g, const uint8_t *volume, double *loss, float *grad)
{
	return cnn_guarded(c, [&] {
		const double l = c->model->gradients(cnn_grid(g), volume, false, grad);
		if (loss) {
			*loss = l;
		}
	});
}

int cgp_cnn_train_step(cgp_cnn c, cgp_graph g, const uint8_t *volume, float learning_rate, double *loss)
{
	return cnn_guarded(c, [&] {
		const double l = c->model->train_step(cnn_grid(g), volume, false, learning_rate);
		if (loss) {
			*loss = l;
		}
	});
}

int cgp_max_flow_certified(cgp_graph g, int64_t certificate, int64_t *flow, cgp_stats *stats, int32_t *stopped_early)
{
	return guarded(g, [&] {
		if (certificate <= 0) {
			cgp::fail(CGP_ERR_INVALID, "certificate must be positive");
		}
		cgp::GridSolver &grid = single_grid(g);
		grid.set_profiling(g->profiling);
		grid.set_certificate(certificate);
		int64_t f = 0;
		try {
			f = grid.solve(g->policy, stats, g->trace);
		} catch (...) {
			grid.set_certificate(0);
			throw;
		}
		grid.set_certificate(0);
		if (flow) {
			*flow = f;
		}
		if (stopped_early) {
			*stopped_early = grid.stopped_early() ? 1 : 0;
		}
	});
}

int cgp_grid_max_flow_noalloc(int32_t nx, int32_t ny, int32_t nz, int32_t conn, const int32_t *cap_source,
                              const int32_t *cap_sink, const int32_t *cap_n, const int32_t *cap_n_back,
                              uint8_t *source_set, int64_t *flow, void *stream)
{
	return guarded_global([&] {
		if (!flow) {
			cgp::fail(CGP_ERR_INVALID, "flow pointer is null");
		}
		*flow = cgp::noalloc_max_flow(nx, ny, nz, conn, cap_source, cap_sink, cap_n, cap_n_back, source_set,
		                              static_cast<hipStream_t>(stream));
	});
}

int cgp_set_grid_batch(cgp_graph g, int32_t count, int32_t conn, const cgp_grid_instance *grids,
                       int32_t device_pointers)
{
	return guarded(g, [&] {
		if (count <= 0 || !grids) {
			cgp::fail(CGP_ERR_INVALID, "a batch needs at least one grid");
		}
		std::vector<cgp::GridSpec> specs((size_t)count);
		for (int32_t i = 0; i < count; ++i) {
			const cgp_grid_instance &in = grids[i];
			specs[(size_t)i] = cgp::GridSpec{in.nx, in.ny, in.nz, in.cap_source, in.cap_sink, in.cap_n, in.cap_n_back};
		}
		g->csr.reset();
		if (!g->grid) {
			g->grid.reset(new cgp::GridSolver(g->stream));
		}
		g->mode = cgp_graph_s::Mode::Grid;
		g->grid->set_batch(conn, specs.data(), count, device_pointers != 0);
	});
}

int cgp_batch_size(cgp_graph g, int32_t *count)
{
	return guarded(g, [&] {
		if (!count) {
			cgp::fail(CGP_ERR_INVALID, "output pointer is null");
		}
		*count = grid_of(g).instance_count();
	});
}

int cgp_solve_batch(cgp_graph g, int64_t *flows, cgp_stats *stats)
{
	return guarded(g, [&] {
		cgp::GridSolver &grid = grid_of(g);
		grid.set_profiling(g->profiling);
		grid.solve(g->policy, stats, g->trace);
		if (flows) {
			const std::vector<int64_t> &f = grid.flows();
			std::memcpy(flows, f.data(), f.size() * sizeof(int64_t));
		}
	});
}

int cgp_batch_source_set(cgp_graph g, int32_t index, uint8_t *out)
{
	return guarded(g, [&] {
		if (!out) {
			cgp::fail(CGP_ERR_INVALID, "output pointer is null");
		}
		grid_of(g).source_side(index, out);
	});
}

int cgp_batch_fill_segments(cgp_graph g, int32_t index, uint8_t *out, uint8_t mask, uint8_t background)
{
	return guarded(g, [&] {
		if (!out) {
			cgp::fail(CGP_ERR_INVALID, "output pointer is null");
		}
		grid_of(g).sink_side(index, out, mask, background);
	});
}

int cgp_set_policy(cgp_graph g, const cgp_policy *policy)
{
	return guarded(g, [&] {
		if (!policy) {
			g->policy = cgp::default_policy();
			return;
		}
		if (policy->minimal_residual < 0.0f) {
			cgp::fail(CGP_ERR_INVALID, "minimal_residual must be non-negative");
		}
		if (policy->tile_x < 0 || policy->tile_y < 0 || policy->tile_z < 0 || policy->discharges_per_relabel < 0 ||
		    policy->max_tile_iterations < 0 || policy->max_rounds < 0) {
			cgp::fail(CGP_ERR_INVALID, "policy fields must be non-negative");
		}
		g->policy = *policy;
	});
}

int cgp_get_policy(cgp_graph g, cgp_policy *policy)
{
	if (!g || !policy) {
		return CGP_ERR_INVALID;
	}
	*policy = g->policy;
	return CGP_OK;
}

int cgp_set_trace(cgp_graph g, cgp_trace_fn fn, void *user)
{
	if (!g) {
		return CGP_ERR_INVALID;
	}
	g->trace.fn = fn;
	g->trace.user = user;
	return CGP_OK;
}

int cgp_max_flow(cgp_graph g, int64_t *flow, cgp_stats *stats)
{
	return guarded(g, [&] {
		int64_t f = 0;
		switch (g->mode) {
		case cgp_graph_s::Mode::Generic:
			f = g->csr->solve(g->policy, stats, g->trace);
			break;
		case cgp_graph_s::Mode::Grid:
			g->grid->set_profiling(g->profiling);
			f = g->grid->solve(g->policy, stats, g->trace);
			break;
		default:
			cgp::fail(CGP_ERR_STATE, "no graph set");
		}
		if (flow) {
			*flow = f;
		}
	});
}

int cgp_fill_segments(cgp_graph g, uint8_t *out, uint8_t mask, uint8_t background)
{
	return guarded(g, [&] {
		if (!out) {
			cgp::fail(CGP_ERR_INVALID, "output pointer is null");
		}
		if (g->mode == cgp_graph_s::Mode::Grid) {
			single_grid(g).sink_side(0, out, mask, background);
		} else {
			generic(g).sink_side(out, mask, background);
		}
	});
}

int cgp_source_set(cgp_graph g, uint8_t *out)
{
	return guarded(g, [&] {
		if (!out) {
			cgp::fail(CGP_ERR_INVALID, "output pointer is null");
		}
		if (g->mode == cgp_graph_s::Mode::Grid) {
			single_grid(g).source_side(0, out);
		} else {
			generic(g).source_side(out);
		}
	});
}

int cgp_labels(cgp_graph g, int32_t *out)
{
	return guarded(g, [&] {
		if (!out) {
			cgp::fail(CGP_ERR_INVALID, "output pointer is null");
		}
		if (g->mode == cgp_graph_s::Mode::Grid) {
			single_grid(g).labels(0, out);
		} else {
			generic(g).labels(out);
		}
	});
}

int cgp_set_profiling(cgp_graph g, int32_t enable)
{
	if (!g) {
		return CGP_ERR_INVALID;
	}
	g->profiling = enable != 0;
	return CGP_OK;
}

int cgp_get_profile(cgp_graph g, cgp_kernel_profile *out, int32_t capacity, int32_t *count)
{
	return guarded(g, [&] {
		if (capacity < 0 || (capacity > 0 && !out)) {
			cgp::fail(CGP_ERR_INVALID, "invalid output array");
		}
		int n = 0;
		if (g->mode == cgp_graph_s::Mode::Grid && g->grid) {
			n = g->grid->profile(out, capacity);
		}
		if (count) {
			*count = n;
		}
	});
}

}  // extern "C"
