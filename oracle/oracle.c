/*
 * CPU oracle for the push-relabel graph cut -- TEST INFRASTRUCTURE ONLY.
 *
 * Only tests/, __graft_entry__.smoke() and the cpu_baseline leg of bench.py may
 * load this library, and only as the checker or as the timed CPU baseline. The
 * product (cugip_amd/, include/) never calls it.
 *
 * Network: vertices 0..n-1 plus terminals s and t; s->v capacity cs[v], v->t
 * capacity ct[v]; edge i joins ea[i] and eb[i] with capacity fwd[i] (ea->eb) and
 * bwd[i] (eb->ea). All capacities int32 >= 0, all flow arithmetic int64.
 * Outputs: the max-flow value, S_min (vertices reachable from s in the residual
 * graph of a maximum flow) and T_min (vertices that reach t in it), as 0/1 bytes.
 *
 * Functions:
 *   orc_refpr      restatement of the reference's algorithm (global BFS relabel from
 *                  the sink frontier, then one top-down level push sweep, until a sweep
 *                  pushes nothing): MinCut::compute,
 *                  include/cugip/advanced_operations/detail/graph_cut_implementation.hpp:122-198;
 *                  Relabel::compute / initBFSKernel / propagateFromVertex,
 *                  detail/graph_cut_relabeling.hpp:47-112, 1004-1029, 1139-1164;
 *                  Push::compute / pushImplementation / tryPullPush,
 *                  detail/graph_cut_push.hpp:96-186, 461-482;
 *                  source preflow, detail/graph_cut_push.hpp:9-24.
 *                  Deliberate deviations: exact integers instead of float; finite sink
 *                  t-links (level 1 is "residual sink capacity > 0" and level-1 vertices
 *                  push min(excess, residual) into t, where the reference drains all
 *                  excess, graph_cut_push.hpp:26-41); every vertex's adjacency is used
 *                  (no neighbors[n] sentinel, graph_cut.hpp:255-270). S_min is the
 *                  residual closure of s and of every vertex left with excess (a maximum
 *                  preflow), T_min the vertices with a finite final label
 *                  (fill_segments / SetMask, graph_cut.hpp:304-333).
 *   orc_dinic      textbook Dinic max flow on the explicit s-t network; S_min by BFS
 *                  from s, T_min by reverse BFS from t. Independent of the above.
 *   orc_bk         Boykov-Kolmogorov (IEEE TPAMI 26(9), 2004): search trees from s and
 *                  t, growth, augmentation, adoption with the distance/timestamp
 *                  heuristic. Stand-in for the reference's CPU comparator (BK301
 *                  Graph::maxflow, benchmarks/graph_cut/boykov_kolmogorov.cpp:19-83),
 *                  which is not shipped with the reference; timed as the CPU baseline.
 *   orc_bruteforce enumeration of all 2^n source sides (n <= 24): min cut value, S_min as
 *                  the intersection and T_min as the complement of the union of all
 *                  minimum source sides.
 *   orc_grid_edges edge list of a 6-/26-connected grid given SoA n-link planes, in the
 *                  voxel numbering and half-neighbourhood order of the reference's grid
 *                  construction (benchmarks/graph_cut/graph_cut.cpp:32-98,
 *                  cugip/neighborhood.hpp:51-91, 207-245).
 */
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

#define ORC_INF_LABEL 0x3fffffff

/* ------------------------------------------------------------------------------------------ */
/* residual network on CSR arcs                                                                */

typedef struct {
	int32_t n;
	int64_t narcs;
	int64_t *first;   /* n + 1 */
	int32_t *head;
	int64_t *mate;
	int64_t *res;
} net_t;

static void net_free(net_t *g)
{
	free(g->first);
	free(g->head);
	free(g->mate);
	free(g->res);
	memset(g, 0, sizeof(*g));
}

/* Build arcs for vertices 0..n-1 from the edge list; loops and edges with both capacities 0
   are dropped (they carry no flow). */
static int net_build(net_t *g, int32_t n, int64_t m, const int32_t *ea, const int32_t *eb, const int32_t *fwd,
                     const int32_t *bwd)
{
	memset(g, 0, sizeof(*g));
	g->n = n;
	g->first = (int64_t *)calloc((size_t)n + 1, sizeof(int64_t));
	if (!g->first) {
		return -1;
	}
	for (int64_t i = 0; i < m; ++i) {
		if (ea[i] == eb[i] || (fwd[i] == 0 && bwd[i] == 0)) {
			continue;
		}
		g->first[ea[i] + 1]++;
		g->first[eb[i] + 1]++;
	}
	for (int32_t v = 0; v < n; ++v) {
		g->first[v + 1] += g->first[v];
	}
	g->narcs = g->first[n];
	g->head = (int32_t *)malloc(sizeof(int32_t) * (size_t)(g->narcs ? g->narcs : 1));
	g->mate = (int64_t *)malloc(sizeof(int64_t) * (size_t)(g->narcs ? g->narcs : 1));
	g->res = (int64_t *)malloc(sizeof(int64_t) * (size_t)(g->narcs ? g->narcs : 1));
	int64_t *pos = (int64_t *)malloc(sizeof(int64_t) * ((size_t)n + 1));
	if (!g->head || !g->mate || !g->res || !pos) {
		free(pos);
		net_free(g);
		return -1;
	}
	memcpy(pos, g->first, sizeof(int64_t) * (size_t)n);
	for (int64_t i = 0; i < m; ++i) {
		const int32_t a = ea[i], b = eb[i];
		if (a == b || (fwd[i] == 0 && bwd[i] == 0)) {
			continue;
		}
		const int64_t ia = pos[a]++, ib = pos[b]++;
		g->head[ia] = b;
		g->res[ia] = fwd[i];
		g->mate[ia] = ib;
		g->head[ib] = a;
		g->res[ib] = bwd[i];
		g->mate[ib] = ia;
	}
	free(pos);
	return 0;
}

static int validate(int32_t n, int64_t m, const int32_t *ea, const int32_t *eb, const int32_t *fwd, const int32_t *bwd,
                    const int32_t *cs, const int32_t *ct)
{
	if (n <= 0) {
		return -1;
	}
	for (int64_t i = 0; i < m; ++i) {
		if (ea[i] < 0 || ea[i] >= n || eb[i] < 0 || eb[i] >= n || fwd[i] < 0 || bwd[i] < 0) {
			return -1;
		}
	}
	for (int32_t v = 0; v < n; ++v) {
		if (cs[v] < 0 || ct[v] < 0) {
			return -1;
		}
	}
	return 0;
}

/* S_min of a maximum preflow: closure under residual arcs of the vertices with excess (the
   source arcs are saturated, so s itself reaches nothing directly). */
static void closure_from_excess(const net_t *g, const int64_t *excess, uint8_t *out)
{
	int32_t *queue = (int32_t *)malloc(sizeof(int32_t) * (size_t)g->n);
	int64_t qh = 0, qt = 0;
	for (int32_t v = 0; v < g->n; ++v) {
		out[v] = excess[v] > 0;
		if (out[v]) {
			queue[qt++] = v;
		}
	}
	while (qh < qt) {
		const int32_t v = queue[qh++];
		for (int64_t a = g->first[v]; a < g->first[v + 1]; ++a) {
			const int32_t u = g->head[a];
			if (g->res[a] > 0 && !out[u]) {
				out[u] = 1;
				queue[qt++] = u;
			}
		}
	}
	free(queue);
}

/* ------------------------------------------------------------------------------------------ */
/* restatement of the reference's relabel / push loop                                           */

typedef struct {
	int64_t iterations;      /* relabel + push rounds */
	int64_t pushes;
	int64_t max_level;
} orc_refpr_stats;

/* Global relabel: BFS from the sink frontier over reverse residual arcs. Fills h and the queue
   ordered by level; level_end[L] = queue end of level L (level_end[0] = 0). Returns the top level. */
static int64_t refpr_relabel(const net_t *g, const int64_t *rt, int32_t *h, int32_t *queue, int64_t *level_end,
                             int64_t max_levels)
{
	int64_t qt = 0;
	for (int32_t v = 0; v < g->n; ++v) {
		h[v] = ORC_INF_LABEL;
		if (rt[v] > 0) {
			h[v] = 1;
			queue[qt++] = v;
		}
	}
	level_end[0] = 0;
	level_end[1] = qt;
	int64_t level = 1;
	int64_t qh = 0;
	while (qh < qt && level < max_levels - 1) {
		const int64_t end = qt;
		for (; qh < end; ++qh) {
			const int32_t v = queue[qh];
			for (int64_t a = g->first[v]; a < g->first[v + 1]; ++a) {
				const int32_t u = g->head[a];
				if (h[u] == ORC_INF_LABEL && g->res[g->mate[a]] > 0) {
					h[u] = (int32_t)(level + 1);
					queue[qt++] = u;
				}
			}
		}
		if (qt == end) {
			break;
		}
		++level;
		level_end[level] = qt;
	}
	return level;
}

int64_t orc_refpr(int32_t n, int64_t m, const int32_t *ea, const int32_t *eb, const int32_t *fwd, const int32_t *bwd,
                  const int32_t *cs, const int32_t *ct, uint8_t *smin, uint8_t *tmin, int32_t *labels_out,
                  orc_refpr_stats *stats)
{
	if (validate(n, m, ea, eb, fwd, bwd, cs, ct)) {
		return -1;
	}
	net_t g;
	if (net_build(&g, n, m, ea, eb, fwd, bwd)) {
		return -2;
	}
	int64_t *e = (int64_t *)malloc(sizeof(int64_t) * (size_t)n);
	int64_t *rt = (int64_t *)malloc(sizeof(int64_t) * (size_t)n);
	int32_t *h = (int32_t *)malloc(sizeof(int32_t) * (size_t)n);
	int32_t *queue = (int32_t *)malloc(sizeof(int32_t) * (size_t)n);
	int64_t *level_end = (int64_t *)malloc(sizeof(int64_t) * ((size_t)n + 2));
	if (!e || !rt || !h || !queue || !level_end) {
		net_free(&g);
		free(e);
		free(rt);
		free(h);
		free(queue);
		free(level_end);
		return -2;
	}
	/* preflow: saturate every source t-link (pushThroughTLinksFromSourceKernel) */
	for (int32_t v = 0; v < n; ++v) {
		e[v] = cs[v];
		rt[v] = ct[v];
	}
	int64_t flow = 0;
	orc_refpr_stats st = {0, 0, 0};
	for (;;) {
		const int64_t top = refpr_relabel(&g, rt, h, queue, level_end, (int64_t)n + 2);
		if (top > st.max_level) {
			st.max_level = top;
		}
		++st.iterations;
		int pushed = 0;
		for (int64_t level = top; level >= 1; --level) {
			for (int64_t i = level_end[level - 1]; i < level_end[level]; ++i) {
				const int32_t v = queue[i];
				if (e[v] <= 0) {
					continue;
				}
				if (level == 1 && rt[v] > 0) {
					const int64_t f = e[v] < rt[v] ? e[v] : rt[v];
					rt[v] -= f;
					e[v] -= f;
					flow += f;
					pushed = 1;
					++st.pushes;
				}
				for (int64_t a = g.first[v]; a < g.first[v + 1] && e[v] > 0; ++a) {
					const int32_t u = g.head[a];
					if (h[u] < level && h[u] >= 0 && g.res[a] > 0) {
						const int64_t f = e[v] < g.res[a] ? e[v] : g.res[a];
						g.res[a] -= f;
						g.res[g.mate[a]] += f;
						e[v] -= f;
						e[u] += f;
						pushed = 1;
						++st.pushes;
					}
				}
			}
		}
		if (!pushed) {
			break;
		}
	}
	if (smin) {
		closure_from_excess(&g, e, smin);
	}
	for (int32_t v = 0; v < n; ++v) {
		if (tmin) {
			tmin[v] = h[v] < ORC_INF_LABEL;
		}
		if (labels_out) {
			labels_out[v] = h[v];
		}
	}
	if (stats) {
		*stats = st;
	}
	net_free(&g);
	free(e);
	free(rt);
	free(h);
	free(queue);
	free(level_end);
	return flow;
}

/* Level-1 seed count of the reference's initBFSKernel (graph_cut_relabeling.hpp:47-80) on the
   initial residual graph: vertices whose sink t-link is positive. */
int64_t orc_bfs_seed_count(int32_t n, const int32_t *ct)
{
	int64_t c = 0;
	for (int32_t v = 0; v < n; ++v) {
		c += ct[v] > 0;
	}
	return c;
}

/* Labels of the reference's first global relabel (before any push): BFS distances to the
   sink frontier, ORC_INF_LABEL when unreachable. */
int orc_initial_labels(int32_t n, int64_t m, const int32_t *ea, const int32_t *eb, const int32_t *fwd,
                       const int32_t *bwd, const int32_t *ct, int32_t *labels_out)
{
	net_t g;
	if (net_build(&g, n, m, ea, eb, fwd, bwd)) {
		return -2;
	}
	int64_t *rt = (int64_t *)malloc(sizeof(int64_t) * (size_t)n);
	int32_t *queue = (int32_t *)malloc(sizeof(int32_t) * (size_t)n);
	int64_t *level_end = (int64_t *)malloc(sizeof(int64_t) * ((size_t)n + 2));
	for (int32_t v = 0; v < n; ++v) {
		rt[v] = ct[v];
	}
	refpr_relabel(&g, rt, labels_out, queue, level_end, (int64_t)n + 2);
	free(rt);
	free(queue);
	free(level_end);
	net_free(&g);
	return 0;
}

/* ------------------------------------------------------------------------------------------ */
/* Dinic on the explicit network (vertices n = s, n + 1 = t)                                    */

static int64_t dinic_dfs(net_t *g, int32_t v, int32_t t, int64_t pushed, const int32_t *level, int64_t *it,
                         int32_t *stack_v, int64_t *stack_a)
{
	/* iterative DFS with current-arc pointers */
	int64_t depth = 0;
	stack_v[0] = v;
	for (;;) {
		const int32_t x = stack_v[depth];
		if (x == t) {
			int64_t f = pushed;
			for (int64_t k = 0; k < depth; ++k) {
				if (g->res[stack_a[k]] < f) {
					f = g->res[stack_a[k]];
				}
			}
			for (int64_t k = 0; k < depth; ++k) {
				g->res[stack_a[k]] -= f;
				g->res[g->mate[stack_a[k]]] += f;
			}
			return f;
		}
		int advanced = 0;
		for (; it[x] < g->first[x + 1]; ++it[x]) {
			const int64_t a = it[x];
			const int32_t y = g->head[a];
			if (g->res[a] > 0 && level[y] == level[x] + 1) {
				stack_a[depth] = a;
				stack_v[++depth] = y;
				advanced = 1;
				break;
			}
		}
		if (!advanced) {
			if (depth == 0) {
				return 0;
			}
			--depth;
			++it[stack_v[depth]];
		}
	}
}

int64_t orc_dinic(int32_t n, int64_t m, const int32_t *ea, const int32_t *eb, const int32_t *fwd, const int32_t *bwd,
                  const int32_t *cs, const int32_t *ct, uint8_t *smin, uint8_t *tmin)
{
	if (validate(n, m, ea, eb, fwd, bwd, cs, ct)) {
		return -1;
	}
	const int32_t s = n, t = n + 1, nn = n + 2;
	const int64_t mm = m + 2 * (int64_t)n;
	int32_t *A = (int32_t *)malloc(sizeof(int32_t) * (size_t)mm);
	int32_t *B = (int32_t *)malloc(sizeof(int32_t) * (size_t)mm);
	int32_t *F = (int32_t *)malloc(sizeof(int32_t) * (size_t)mm);
	int32_t *R = (int32_t *)malloc(sizeof(int32_t) * (size_t)mm);
	if (!A || !B || !F || !R) {
		free(A);
		free(B);
		free(F);
		free(R);
		return -2;
	}
	memcpy(A, ea, sizeof(int32_t) * (size_t)m);
	memcpy(B, eb, sizeof(int32_t) * (size_t)m);
	memcpy(F, fwd, sizeof(int32_t) * (size_t)m);
	memcpy(R, bwd, sizeof(int32_t) * (size_t)m);
	for (int32_t v = 0; v < n; ++v) {
		A[m + 2 * v] = s;
		B[m + 2 * v] = v;
		F[m + 2 * v] = cs[v];
		R[m + 2 * v] = 0;
		A[m + 2 * v + 1] = v;
		B[m + 2 * v + 1] = t;
		F[m + 2 * v + 1] = ct[v];
		R[m + 2 * v + 1] = 0;
	}
	net_t g;
	const int rc = net_build(&g, nn, mm, A, B, F, R);
	free(A);
	free(B);
	free(F);
	free(R);
	if (rc) {
		return -2;
	}
	int32_t *level = (int32_t *)malloc(sizeof(int32_t) * (size_t)nn);
	int32_t *queue = (int32_t *)malloc(sizeof(int32_t) * (size_t)nn);
	int64_t *it = (int64_t *)malloc(sizeof(int64_t) * (size_t)nn);
	int32_t *stack_v = (int32_t *)malloc(sizeof(int32_t) * ((size_t)nn + 1));
	int64_t *stack_a = (int64_t *)malloc(sizeof(int64_t) * ((size_t)nn + 1));
	int64_t flow = 0;
	for (;;) {
		for (int32_t v = 0; v < nn; ++v) {
			level[v] = -1;
		}
		int64_t qh = 0, qt = 0;
		level[s] = 0;
		queue[qt++] = s;
		while (qh < qt) {
			const int32_t v = queue[qh++];
			for (int64_t a = g.first[v]; a < g.first[v + 1]; ++a) {
				const int32_t u = g.head[a];
				if (g.res[a] > 0 && level[u] < 0) {
					level[u] = level[v] + 1;
					queue[qt++] = u;
				}
			}
		}
		if (level[t] < 0) {
			break;
		}
		for (int32_t v = 0; v < nn; ++v) {
			it[v] = g.first[v];
		}
		for (;;) {
			const int64_t f = dinic_dfs(&g, s, t, INT64_MAX, level, it, stack_v, stack_a);
			if (f == 0) {
				break;
			}
			flow += f;
		}
	}
	/* S_min: BFS from s (level[] of the last BFS holds exactly the s-reachable set) */
	if (smin) {
		for (int32_t v = 0; v < n; ++v) {
			smin[v] = level[v] >= 0;
		}
	}
	/* T_min: reverse BFS from t over arcs u -> v with residual > 0 */
	if (tmin) {
		uint8_t *seen = (uint8_t *)calloc((size_t)nn, 1);
		int64_t qh = 0, qt = 0;
		seen[t] = 1;
		queue[qt++] = t;
		while (qh < qt) {
			const int32_t v = queue[qh++];
			for (int64_t a = g.first[v]; a < g.first[v + 1]; ++a) {
				const int32_t u = g.head[a];
				if (!seen[u] && g.res[g.mate[a]] > 0) {
					seen[u] = 1;
					queue[qt++] = u;
				}
			}
		}
		for (int32_t v = 0; v < n; ++v) {
			tmin[v] = seen[v];
		}
		free(seen);
	}
	free(level);
	free(queue);
	free(it);
	free(stack_v);
	free(stack_a);
	net_free(&g);
	return flow;
}

/* ------------------------------------------------------------------------------------------ */
/* Boykov-Kolmogorov                                                                            */

#define BK_NONE (-1)      /* parent of a free vertex */
#define BK_TERMINAL (-2)  /* parent of a vertex attached directly to its terminal */
#define BK_ORPHAN (-3)
#define BK_DIST_INF 0x7fffffff

typedef struct {
	net_t g;
	int64_t *tr;         /* > 0: residual s->v, < 0: residual v->t (after cancellation) */
	int64_t *parent;     /* arc from the vertex to its parent, or BK_* */
	uint8_t *tree;       /* 0: S, 1: T (meaningful when parent != BK_NONE) */
	int32_t *dist;
	int64_t *ts;
	uint8_t *in_active;
	int32_t *active;     /* circular FIFO of capacity n */
	int64_t ahead, acount;
	int32_t *orphans;    /* FIFO, capacity n */
	int64_t ohead, ocount;
	int64_t time;
} bk_t;

static void bk_activate(bk_t *b, int32_t v)
{
	if (!b->in_active[v]) {
		b->in_active[v] = 1;
		b->active[(b->ahead + b->acount) % b->g.n] = v;
		b->acount++;
	}
}

static int32_t bk_next_active(bk_t *b)
{
	while (b->acount > 0) {
		const int32_t v = b->active[b->ahead];
		b->ahead = (b->ahead + 1) % b->g.n;
		b->acount--;
		b->in_active[v] = 0;
		if (b->parent[v] != BK_NONE) {
			return v;
		}
	}
	return -1;
}

static void bk_orphan(bk_t *b, int32_t v)
{
	b->parent[v] = BK_ORPHAN;
	b->orphans[(b->ohead + b->ocount) % b->g.n] = v;
	b->ocount++;
}

static int64_t bk_augment(bk_t *b, int64_t arc_st)
{
	net_t *g = &b->g;
	/* bottleneck */
	int64_t f = g->res[arc_st];
	int32_t v = g->head[g->mate[arc_st]]; /* tail of arc_st, in S */
	for (;;) {
		const int64_t a = b->parent[v];
		if (a == BK_TERMINAL) {
			break;
		}
		if (g->res[g->mate[a]] < f) {
			f = g->res[g->mate[a]];
		}
		v = g->head[a];
	}
	if (b->tr[v] < f) {
		f = b->tr[v];
	}
	v = g->head[arc_st]; /* in T */
	for (;;) {
		const int64_t a = b->parent[v];
		if (a == BK_TERMINAL) {
			break;
		}
		if (g->res[a] < f) {
			f = g->res[a];
		}
		v = g->head[a];
	}
	if (-b->tr[v] < f) {
		f = -b->tr[v];
	}
	/* augment */
	g->res[arc_st] -= f;
	g->res[g->mate[arc_st]] += f;
	v = g->head[g->mate[arc_st]];
	for (;;) {
		const int64_t a = b->parent[v];
		if (a == BK_TERMINAL) {
			break;
		}
		g->res[a] += f;
		g->res[g->mate[a]] -= f;
		if (g->res[g->mate[a]] == 0) {
			bk_orphan(b, v);
		}
		v = g->head[a];
	}
	b->tr[v] -= f;
	if (b->tr[v] == 0) {
		bk_orphan(b, v);
	}
	v = g->head[arc_st];
	for (;;) {
		const int64_t a = b->parent[v];
		if (a == BK_TERMINAL) {
			break;
		}
		g->res[a] -= f;
		g->res[g->mate[a]] += f;
		if (g->res[a] == 0) {
			bk_orphan(b, v);
		}
		v = g->head[a];
	}
	b->tr[v] += f;
	if (b->tr[v] == 0) {
		bk_orphan(b, v);
	}
	return f;
}

/* find a new parent for orphan v in tree `tr` (0 = S, 1 = T), else free it */
static void bk_adopt(bk_t *b, int32_t v)
{
	net_t *g = &b->g;
	const uint8_t tr = b->tree[v];
	int64_t best_arc = BK_NONE;
	int32_t best_d = BK_DIST_INF;
	for (int64_t a0 = g->first[v]; a0 < g->first[v + 1]; ++a0) {
		/* S: need residual u -> v; T: need residual v -> u */
		const int64_t cap = tr == 0 ? g->res[g->mate[a0]] : g->res[a0];
		if (cap <= 0) {
			continue;
		}
		int32_t u = g->head[a0];
		if (b->parent[u] == BK_NONE || b->tree[u] != tr) {
			continue;
		}
		/* is u rooted at the terminal? walk up, counting steps */
		int32_t d = 0;
		int32_t x = u;
		for (;;) {
			if (b->ts[x] == b->time) {
				d += b->dist[x];
				break;
			}
			const int64_t pa = b->parent[x];
			++d;
			if (pa == BK_TERMINAL) {
				b->ts[x] = b->time;
				b->dist[x] = 1;
				break;
			}
			if (pa == BK_ORPHAN) {
				d = BK_DIST_INF;
				break;
			}
			x = g->head[pa];
		}
		if (d < BK_DIST_INF) {
			if (d < best_d) {
				best_arc = a0;
				best_d = d;
			}
			/* record distances along the verified path */
			for (x = u; b->ts[x] != b->time; x = g->head[b->parent[x]]) {
				b->ts[x] = b->time;
				b->dist[x] = d--;
			}
		}
	}
	if (best_arc != BK_NONE) {
		b->parent[v] = best_arc;
		b->ts[v] = b->time;
		b->dist[v] = best_d + 1;
		return;
	}
	/* no parent: v becomes free; its tree neighbours may grow into it again, and its
	   children become orphans */
	for (int64_t a0 = g->first[v]; a0 < g->first[v + 1]; ++a0) {
		const int32_t u = g->head[a0];
		const int64_t pa = b->parent[u];
		if (pa == BK_NONE || b->tree[u] != tr) {
			continue;
		}
		const int64_t cap = tr == 0 ? g->res[g->mate[a0]] : g->res[a0];
		if (cap > 0) {
			bk_activate(b, u);
		}
		if (pa != BK_TERMINAL && pa != BK_ORPHAN && g->head[pa] == v) {
			bk_orphan(b, u);
		}
	}
	b->parent[v] = BK_NONE;
}

int64_t orc_bk(int32_t n, int64_t m, const int32_t *ea, const int32_t *eb, const int32_t *fwd, const int32_t *bwd,
               const int32_t *cs, const int32_t *ct, uint8_t *smin, uint8_t *tmin, double *maxflow_seconds)
{
	if (validate(n, m, ea, eb, fwd, bwd, cs, ct)) {
		return -1;
	}
	bk_t b;
	memset(&b, 0, sizeof(b));
	if (net_build(&b.g, n, m, ea, eb, fwd, bwd)) {
		return -2;
	}
	b.tr = (int64_t *)malloc(sizeof(int64_t) * (size_t)n);
	b.parent = (int64_t *)malloc(sizeof(int64_t) * (size_t)n);
	b.tree = (uint8_t *)calloc((size_t)n, 1);
	b.dist = (int32_t *)malloc(sizeof(int32_t) * (size_t)n);
	b.ts = (int64_t *)calloc((size_t)n, sizeof(int64_t));
	b.in_active = (uint8_t *)calloc((size_t)n, 1);
	b.active = (int32_t *)malloc(sizeof(int32_t) * (size_t)n);
	b.orphans = (int32_t *)malloc(sizeof(int32_t) * (size_t)n);
	int64_t flow = 0;
	for (int32_t v = 0; v < n; ++v) {
		const int64_t c = cs[v] < ct[v] ? cs[v] : ct[v];
		flow += c;
		b.tr[v] = (int64_t)cs[v] - ct[v];
	}

	struct timespec t0, t1;
	clock_gettime(CLOCK_MONOTONIC, &t0);
	for (int32_t v = 0; v < n; ++v) {
		if (b.tr[v] != 0) {
			b.tree[v] = b.tr[v] > 0 ? 0 : 1;
			b.parent[v] = BK_TERMINAL;
			b.dist[v] = 1;
			b.ts[v] = 0;
			bk_activate(&b, v);
		} else {
			b.parent[v] = BK_NONE;
		}
	}
	net_t *g = &b.g;
	int32_t cur = -1;
	for (;;) {
		if (cur < 0 || b.parent[cur] == BK_NONE) {
			cur = bk_next_active(&b);
			if (cur < 0) {
				break;
			}
		}
		/* growth */
		int64_t path = BK_NONE;
		if (b.tree[cur] == 0) {
			for (int64_t a = g->first[cur]; a < g->first[cur + 1]; ++a) {
				if (g->res[a] <= 0) {
					continue;
				}
				const int32_t u = g->head[a];
				if (b.parent[u] == BK_NONE) {
					b.tree[u] = 0;
					b.parent[u] = g->mate[a];
					b.ts[u] = b.ts[cur];
					b.dist[u] = b.dist[cur] + 1;
					bk_activate(&b, u);
				} else if (b.tree[u] == 1) {
					path = a;
					break;
				} else if (b.ts[u] <= b.ts[cur] && b.dist[u] > b.dist[cur]) {
					b.parent[u] = g->mate[a];
					b.ts[u] = b.ts[cur];
					b.dist[u] = b.dist[cur] + 1;
				}
			}
		} else {
			for (int64_t a = g->first[cur]; a < g->first[cur + 1]; ++a) {
				if (g->res[g->mate[a]] <= 0) {
					continue;
				}
				const int32_t u = g->head[a];
				if (b.parent[u] == BK_NONE) {
					b.tree[u] = 1;
					b.parent[u] = g->mate[a];
					b.ts[u] = b.ts[cur];
					b.dist[u] = b.dist[cur] + 1;
					bk_activate(&b, u);
				} else if (b.tree[u] == 0) {
					path = g->mate[a]; /* u -> cur */
					break;
				} else if (b.ts[u] <= b.ts[cur] && b.dist[u] > b.dist[cur]) {
					b.parent[u] = g->mate[a];
					b.ts[u] = b.ts[cur];
					b.dist[u] = b.dist[cur] + 1;
				}
			}
		}
		if (path == BK_NONE) {
			cur = -1;
			continue;
		}
		/* cur stays the current vertex: growth resumes from it after the adoption */
		flow += bk_augment(&b, path);
		b.time++;
		/* adoption */
		while (b.ocount > 0) {
			const int32_t v = b.orphans[b.ohead];
			b.ohead = (b.ohead + 1) % n;
			b.ocount--;
			bk_adopt(&b, v);
		}
	}
	clock_gettime(CLOCK_MONOTONIC, &t1);
	if (maxflow_seconds) {
		*maxflow_seconds = (double)(t1.tv_sec - t0.tv_sec) + 1e-9 * (double)(t1.tv_nsec - t0.tv_nsec);
	}
	/* S_min: residual closure of the vertices with residual source capacity */
	if (smin) {
		int32_t *queue = (int32_t *)malloc(sizeof(int32_t) * (size_t)n);
		int64_t qh = 0, qt = 0;
		for (int32_t v = 0; v < n; ++v) {
			smin[v] = b.tr[v] > 0;
			if (smin[v]) {
				queue[qt++] = v;
			}
		}
		while (qh < qt) {
			const int32_t v = queue[qh++];
			for (int64_t a = g->first[v]; a < g->first[v + 1]; ++a) {
				const int32_t u = g->head[a];
				if (g->res[a] > 0 && !smin[u]) {
					smin[u] = 1;
					queue[qt++] = u;
				}
			}
		}
		free(queue);
	}
	if (tmin) {
		int32_t *queue = (int32_t *)malloc(sizeof(int32_t) * (size_t)n);
		int64_t qh = 0, qt = 0;
		for (int32_t v = 0; v < n; ++v) {
			tmin[v] = b.tr[v] < 0;
			if (tmin[v]) {
				queue[qt++] = v;
			}
		}
		while (qh < qt) {
			const int32_t v = queue[qh++];
			for (int64_t a = g->first[v]; a < g->first[v + 1]; ++a) {
				const int32_t u = g->head[a];
				if (g->res[g->mate[a]] > 0 && !tmin[u]) {
					tmin[u] = 1;
					queue[qt++] = u;
				}
			}
		}
		free(queue);
	}
	free(b.tr);
	free(b.parent);
	free(b.tree);
	free(b.dist);
	free(b.ts);
	free(b.in_active);
	free(b.active);
	free(b.orphans);
	net_free(&b.g);
	return flow;
}

/* ------------------------------------------------------------------------------------------ */
/* brute force                                                                                  */

int64_t orc_bruteforce(int32_t n, int64_t m, const int32_t *ea, const int32_t *eb, const int32_t *fwd,
                       const int32_t *bwd, const int32_t *cs, const int32_t *ct, uint8_t *smin, uint8_t *tmin)
{
	if (n > 24 || validate(n, m, ea, eb, fwd, bwd, cs, ct)) {
		return -1;
	}
	int64_t best = INT64_MAX;
	uint32_t inter = 0, uni = 0;
	const uint32_t total = 1u << n;
	for (uint32_t X = 0; X < total; ++X) {
		int64_t c = 0;
		for (int32_t v = 0; v < n; ++v) {
			c += (X >> v) & 1u ? ct[v] : cs[v];
		}
		for (int64_t i = 0; i < m; ++i) {
			const uint32_t ia = (X >> ea[i]) & 1u, ib = (X >> eb[i]) & 1u;
			if (ia && !ib) {
				c += fwd[i];
			} else if (ib && !ia) {
				c += bwd[i];
			}
		}
		if (c < best) {
			best = c;
			inter = X;
			uni = X;
		} else if (c == best) {
			inter &= X;
			uni |= X;
		}
	}
	for (int32_t v = 0; v < n; ++v) {
		if (smin) {
			smin[v] = (inter >> v) & 1u;
		}
		if (tmin) {
			tmin[v] = !((uni >> v) & 1u);
		}
	}
	return best;
}

/* ------------------------------------------------------------------------------------------ */
/* grid edge list                                                                               */

static const int8_t HALF6[3][3] = {{-1, 0, 0}, {0, -1, 0}, {0, 0, -1}};

static void half_offset(int32_t conn, int32_t j, int32_t *dx, int32_t *dy, int32_t *dz)
{
	if (conn == 6) {
		*dx = HALF6[j][0];
		*dy = HALF6[j][1];
		*dz = HALF6[j][2];
	} else {
		/* MooreNeighborhood<3>::offset(1 + j) */
		*dx = j % 3 - 1;
		*dy = (j / 3) % 3 - 1;
		*dz = j / 9 - 1;
	}
}

/* Writes the edges (v, v + o_j) with fwd = cap_n[j][v], bwd = cap_nb[j][v] (or cap_n when
   cap_nb is NULL); edges with both capacities 0 or leaving the volume are skipped. Output
   arrays need room for (conn/2) * nx*ny*nz entries. Returns the edge count. */
int64_t orc_grid_edges(int32_t nx, int32_t ny, int32_t nz, int32_t conn, const int32_t *cap_n, const int32_t *cap_nb,
                       int32_t *ea, int32_t *eb, int32_t *fwd, int32_t *bwd)
{
	if (conn != 6 && conn != 26) {
		return -1;
	}
	const int64_t vox = (int64_t)nx * ny * nz;
	int32_t off[13][3];
	for (int32_t j = 0; j < conn / 2; ++j) {
		half_offset(conn, j, &off[j][0], &off[j][1], &off[j][2]);
	}
	/* voxel-major (the half neighbourhood of each voxel in turn): the arcs of one vertex are then
	   written close together when a network is built from the list */
	int64_t m = 0;
	for (int32_t z = 0; z < nz; ++z) {
		for (int32_t y = 0; y < ny; ++y) {
			for (int32_t x = 0; x < nx; ++x) {
				const int64_t v = ((int64_t)z * ny + y) * nx + x;
				for (int32_t j = 0; j < conn / 2; ++j) {
					const int32_t ux = x + off[j][0], uy = y + off[j][1], uz = z + off[j][2];
					if (ux < 0 || ux >= nx || uy < 0 || uy >= ny || uz < 0 || uz >= nz) {
						continue;
					}
					const int64_t u = ((int64_t)uz * ny + uy) * nx + ux;
					const int32_t f = cap_n[j * vox + v];
					const int32_t b = cap_nb ? cap_nb[j * vox + v] : f;
					if (f == 0 && b == 0) {
						continue;
					}
					ea[m] = (int32_t)v;
					eb[m] = (int32_t)u;
					fwd[m] = f;
					bwd[m] = b;
					++m;
				}
			}
		}
	}
	return m;
}
